// stz_api.hip -- the C-ABI (include/stz.h): context, state I/O, and the host
// orchestration of one main-loop iteration of the reference (STZ/bmg.c:138-411).
#include "stz_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

namespace {

// Init (STZ/init.c:6-75) with the compile-time constants as parameters
void derive(stz_ctx *c)
{
    const stz_params &p = c->p;
    StzConsts &k = c->k;
    k.E = p.E;
    k.nu = p.nu;
    if (p.double_nu) k.nu *= 2;                                              // :26
    k.mu = k.E / (2 + 2 * k.nu);                                             // :27
    if (p.dim == 2) k.lambda = k.E * k.nu / (1 - k.nu * k.nu);               // :44
    else k.lambda = k.E * k.nu / ((1 + k.nu) * (1 - 2 * k.nu));
    k.alpha = (k.lambda + k.mu) / (k.lambda + 2.0 * k.mu);                   // elastic.c:786
    k.kT = p.temperature * STZ_BOLZ * STZ_J_IN_EV;                           // bmg.h:73
    k.dt = p.strainincrement / p.strainrate;                                 // :48
    k.tTemp = 1 / (p.trial_frequency * std::exp(-p.temp_soft_barrier / k.kT)); // :49
    k.voxel = p.lx * p.ly * p.lz;                                            // bmg.h:90
    k.ratepre = k.voxel * p.trial_frequency;                                 // evolution.c:125
    k.permsoft = p.perm_soft;
    k.tempsoft = p.temp_soft;
    k.permcap = p.perm_cap;
    k.thr = p.athermal_threshold;
    k.half_gamma0 = 0.5 * p.gamma0;                                          // elastic.c:127
}

// isotropicinv (STZ/elastic.c:522-531), the form used by the loading
void isotropicinv(const stz_ctx *c, const double *st, double *e)
{
    const double lam = c->k.lambda, mu = c->k.mu;
    if (c->C == 3) {
        double tra = -lam / 2 / mu / (2 * lam + 2 * mu) * (st[0] + st[1]);
        e[0] = tra + st[0] / 2 / mu;
        e[1] = tra + st[1] / 2 / mu;
        e[2] = st[2] / 2 / mu;
    } else {
        double tra = -lam / 2 / mu / (3 * lam + 2 * mu) * ((st[0] + st[1]) + st[2]);
        for (int i = 0; i < 3; ++i) e[i] = tra + st[i] / 2 / mu;
        for (int i = 3; i < 6; ++i) e[i] = st[i] / 2 / mu;
    }
}

template <typename R>
__global__ void k_fill(R *p, long long n, R v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename R>
__global__ void k_add(R *p, long long n, double d)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = p[i] + (R)d;
}

// plain 16-byte-per-lane copy: the bandwidth and byte-counter reference of stz_copy_probe
__global__ __launch_bounds__(256) void k_copy16(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// waits for the stream by polling, which returns sooner than a blocking wait
int spin_sync(stz_ctx *ctx)
{
    for (;;) {
        const hipError_t e = hipStreamQuery(ctx->stream);
        if (e == hipSuccess) return STZ_OK;
        if (e != hipErrorNotReady) {
            ctx->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
            return STZ_E_HIP;
        }
    }
}

float elapsed(stz_ctx *ctx, int a, int b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->tev[a], ctx->tev[b]) != hipSuccess) return 0.f;
    return ms;
}

void acc_rates(stz_ctx *ctx)
{
    if (!ctx->timing) return;
    ctx->tms[0] += elapsed(ctx, 8, 9);
    ctx->tn[0]++;
    ctx->tms[5] += elapsed(ctx, 9, 10);
    ctx->tn[5]++;
}

void acc_select(stz_ctx *ctx)
{
    if (!ctx->timing) return;
    ctx->tms[2] += elapsed(ctx, 11, 12);
    ctx->tn[2]++;
}

void acc_solve(stz_ctx *ctx)
{
    if (!ctx->timing) return;
    for (int i = 0; i < 6; ++i) {
        ctx->sms[i] += elapsed(ctx, i, i + 1);
        ctx->sn[i]++;
    }
    ctx->tms[1] += elapsed(ctx, 0, 6);
    ctx->tn[1]++;
}

int dalloc(stz_ctx *ctx, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    STZ_HIP(hipMalloc(p, bytes));
    STZ_HIP(hipMemsetAsync(*p, 0, bytes, ctx->stream));
    return STZ_OK;
}

size_t field_count(const stz_ctx *c, int f)
{
    const size_t N = (size_t)c->N;
    switch (f) {
    case STZ_F_EPS0: case STZ_F_SIGMA: return (size_t)c->C * N;
    case STZ_F_Q0: return (size_t)c->M * N;
    case STZ_F_DEPS: return (size_t)c->M * c->S * N;
    case STZ_F_DEPS_LAST: return (size_t)c->S * N;
    case STZ_F_PERM: case STZ_F_TEMP: case STZ_F_TLAST: case STZ_F_MINQ: case STZ_F_TLOCAL: case STZ_F_MRO: return N;
    default: return 0;
    }
}

// copy doubles into a device array of the storage precision
int put_real(stz_ctx *ctx, void *dst, const double *src, size_t n, size_t offset = 0)
{
    if (ctx->rsz == 8) {
        STZ_HIP(hipMemcpyAsync((double *)dst + offset, src, n * 8, hipMemcpyHostToDevice, ctx->stream));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
        STZ_HIP(hipMemcpyAsync((float *)dst + offset, tmp.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
        STZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    return STZ_OK;
}

int get_real(stz_ctx *ctx, const void *src, double *dst, size_t n)
{
    if (ctx->rsz == 8) {
        STZ_HIP(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        STZ_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(n);
        STZ_HIP(hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        STZ_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < n; ++i) dst[i] = (double)tmp[i];
    }
    return STZ_OK;
}

// applies a pending FP64 uniform stress increment to the field now
int flush_pending(stz_ctx *ctx)
{
    if (ctx->pend_comp < 0) return STZ_OK;
    const long long n = ctx->N;
    double *p = (double *)ctx->sigma + (long long)ctx->pend_comp * n;
    hipLaunchKernelGGL((k_add<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n, ctx->pend_delta);
    STZ_CHECK_LAUNCH();
    ctx->pend_comp = -1;
    ctx->pend_delta = 0.0;
    return STZ_OK;
}

// marks every cached Q0*exp(-D*S) stale (after an upload of any of its inputs)
int invalidate_cache(stz_ctx *ctx)
{
    const long long n = ctx->N;
    const dim3 g((unsigned)((n + 255) / 256));
    if (ctx->rsz == 8)
        hipLaunchKernelGGL((k_fill<double>), g, dim3(256), 0, ctx->stream, (double *)ctx->scache, n, std::nan(""));
    else
        hipLaunchKernelGGL((k_fill<float>), g, dim3(256), 0, ctx->stream, (float *)ctx->scache, n, std::nanf(""));
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int ensure_dnew(stz_ctx *ctx, size_t count)
{
    if (count <= ctx->dnew_cap) return STZ_OK;
    if (ctx->dnew) STZ_HIP(hipFree(ctx->dnew));
    size_t cap = count < 4096 ? 4096 : count;
    STZ_HIP(hipMalloc(&ctx->dnew, cap * sizeof(double)));
    ctx->dnew_cap = cap;
    return STZ_OK;
}

int upload_soff(stz_ctx *ctx)
{
    STZ_HIP(hipMemcpyAsync(ctx->res + 16, ctx->soff, 6 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    return STZ_OK;
}

void draw_voxel_deps(stz_ctx *ctx, double *out)
{
    // VoxelShearTransform redraw, STZ/elastic.c:508-512: mode by mode, components in
    // stored order, each 0.5*GAMMA0 * gsl_ran_gaussian(r, 1.0)
    const int MS = ctx->M * ctx->S;
    for (int i = 0; i < MS; ++i) out[i] = ctx->k.half_gamma0 * ctx->rng.gaussian(1.0);
}

int solve(stz_ctx *ctx)
{
    ctx->pend_comp = -1;
    ctx->pend_delta = 0.0;
    for (int c = 0; c < 6; ++c) ctx->soff[c] = 0.0;
    STZ_HIP(hipMemsetAsync(ctx->res + 15, 0, sizeof(double), ctx->stream));
    // per-voxel moduli: the fixed-point loop of homogeneous solves, else one solve
    int rc = ctx->lamx ? stz_inhom_run(ctx) : stz_solve_launch(ctx, 0);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->hres, ctx->res, 16 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    int nan;
    std::memcpy(&nan, ctx->hres + 15, sizeof(int));
    if (nan) {
        ctx->err = "NaN in the stress solve";
        return STZ_E_NAN;
    }
    for (int c = 0; c < ctx->C; ++c) ctx->stress[c] = ctx->hres[c];
    ctx->nsolve++;
    stz_local_mark_solved(ctx);
    acc_solve(ctx);
    return STZ_OK;
}

// HeatConducting(hdt) (STZ/elastic.c:549-582) of the local temperature; a NaN is reported
// as STZ_E_NAN, where the reference aborts with -4
int heat_step(stz_ctx *ctx, double hdt)
{
    STZ_HIP(hipMemsetAsync(ctx->res + 15, 0, sizeof(double), ctx->stream));
    int rc = stz_heat_launch(ctx, hdt);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->hres + 15, ctx->res + 15, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    int nan;
    std::memcpy(&nan, ctx->hres + 15, sizeof(int));
    if (nan) {
        ctx->err = "NaN in the heat conduction";
        return STZ_E_NAN;
    }
    return STZ_OK;
}

int add_uniform(stz_ctx *ctx, int comp, double delta)
{
    if (ctx->prec == STZ_FP64) {
        if (ctx->pend_comp >= 0) {
            int rc = flush_pending(ctx);
            if (rc) return rc;
        }
        ctx->pend_comp = comp;
        ctx->pend_delta = delta;
        ctx->usum[comp] += delta;     // in the field once applied
    } else {
        ctx->soff[comp] += delta;
    }
    return STZ_OK;
}

int rates(stz_ctx *ctx, long long *nneg, double *R)
{
    int rc = stz_rates_launch(ctx, 0);
    if (rc) return rc;
    ctx->pend_comp = -1;     // consumed by the rate pass
    ctx->pend_delta = 0.0;
    STZ_HIP(hipMemcpyAsync(ctx->hres + 8, ctx->res + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx->rate_total = ctx->hres[8];
    acc_rates(ctx);
    *R = ctx->hres[8];
    *nneg = (long long)ctx->hres[9];
    return STZ_OK;
}

int select_event(stz_ctx *ctx, double dice, int *site, int *mode, double *q)
{
    if (!(dice > 0.0)) {
        ctx->err = "KMC failed to locate an activation mode";
        return STZ_E_SELECT;
    }
    int rc = stz_select_launch(ctx, dice * ctx->rate_total, 0);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->hres + 10, ctx->res + 10, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    acc_select(ctx);
    if (ctx->hres[10] < 0) {
        ctx->err = "KMC failed to locate an activation mode";
        return STZ_E_SELECT;
    }
    *site = (int)ctx->hres[10];
    *mode = (int)ctx->hres[11];
    *q = ctx->hres[12];
    return STZ_OK;
}

int apply_event(stz_ctx *ctx, int site, int mode, const double *dnew)
{
    const size_t ms = (size_t)ctx->M * ctx->S;
    int rc = ensure_dnew(ctx, ms);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->dnew, dnew, ms * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = stz_apply_launch(ctx, site, mode, 0);
    if (rc) return rc;
    ctx->ntrans++;
    return STZ_OK;
}

int athermal_scan(stz_ctx *ctx, int *n)
{
    int rc = stz_athermal_launch(ctx);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->hres + 13, ctx->res + 13, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx->n_athermal = (int)ctx->hres[13];
    *n = ctx->n_athermal;
    return STZ_OK;
}

int athermal_apply(stz_ctx *ctx, const double *dnew, double *soc)
{
    const int n = ctx->n_athermal;
    const size_t cnt = (size_t)n * ctx->M * ctx->S;
    int rc = ensure_dnew(ctx, cnt);
    if (rc) return rc;
    if (n > 0) STZ_HIP(hipMemcpyAsync(ctx->dnew, dnew, cnt * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = upload_soff(ctx);
    if (rc) return rc;
    rc = stz_athermal_apply_launch(ctx, n);
    if (rc) return rc;
    const int nb = (n + 255) / 256;
    std::vector<double> part(nb > 0 ? nb : 1);
    if (nb > 0) STZ_HIP(hipMemcpyAsync(part.data(), ctx->socp, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    if (soc) *soc = s;
    ctx->ntrans += n;
    return STZ_OK;
}

// athermal error subtraction, STZ/evolution.c:91-97 (reference) or loading-consistent
void athermal_correction(stz_ctx *ctx)
{
    double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
    if (ctx->p.athermal_correction == 0) {
        for (int c = 1; c < ctx->C; ++c) ts[c] = ctx->stress[c];
    } else {
        for (int c = 0; c < ctx->C; ++c) ts[c] = (c == ctx->lc) ? 0.0 : ctx->stress[c];
    }
    isotropicinv(ctx, ts, te);
    for (int c = 0; c < ctx->C; ++c) ctx->epsavg[c] -= te[c];
}

// KMC loading after a thermal event (STZ/evolution.c:281-315)
int load_thermal(stz_ctx *ctx, double t)
{
    const double sr = ctx->p.strainrate, mu = ctx->k.mu;
    const int C = ctx->C, lc = ctx->lc;
    double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
    if (ctx->p.loading == STZ_SHEAR) {
        for (int c = 0; c < C; ++c) ts[c] = (c == lc) ? 0 : ctx->stress[c];
        isotropicinv(ctx, ts, te);
        for (int c = 0; c < C; ++c)
            if (c != lc) ctx->epsavg[c] += -te[c];
        ctx->epsavg[lc] += sr * t;
        ctx->stress[lc] += 2 * mu * sr * t;
        return add_uniform(ctx, lc, 2 * mu * sr * t);
    }
    const double nu = ctx->k.nu, E = ctx->k.E;
    for (int c = 1; c < C; ++c) ts[c] = ctx->stress[c];
    isotropicinv(ctx, ts, te);
    ctx->epsavg[0] += (sr * t * 1 - te[0]);
    ctx->epsavg[1] += (sr * t * (-nu) - te[1]);
    if (C == 3) {
        ctx->epsavg[2] += (-te[2]);
    } else {
        ctx->epsavg[2] += (sr * t * (-nu) - te[2]);
        for (int c = 3; c < 6; ++c) ctx->epsavg[c] += (-te[c]);
    }
    ctx->stress[0] += E * sr * t;
    return add_uniform(ctx, 0, E * sr * t);
}

// pure elastic increment (STZ/evolution.c:321-358)
int load_elastic(stz_ctx *ctx)
{
    const double sr = ctx->p.strainrate, inc = ctx->p.strainincrement, mu = ctx->k.mu;
    const int C = ctx->C, lc = ctx->lc;
    if (ctx->p.loading == STZ_SHEAR) {
        ctx->epsavg[lc] += sr * 1;
        ctx->stress[lc] += 2 * mu * inc;
        return add_uniform(ctx, lc, 2 * mu * inc);
    }
    const double nu = ctx->k.nu, E = ctx->k.E;
    ctx->epsavg[0] += (inc * 1);
    ctx->epsavg[1] += (inc * (-nu));
    if (C == 3) {
        ctx->epsavg[2] += (inc * 0.0);
    } else {
        ctx->epsavg[2] += (inc * (-nu));
        for (int c = 3; c < 6; ++c) ctx->epsavg[c] += (inc * 0.0);
    }
    ctx->stress[0] += E * inc;
    return add_uniform(ctx, 0, E * inc);
}

bool terminated(const stz_ctx *ctx)
{
    if (std::fabs(ctx->epsavg[ctx->lc]) > ctx->endstrain) return true;                 // bmg.c:383,397
    if (ctx->p.loading == STZ_TENSION && ctx->stress[0] < -0.5) return true;            // bmg.c:388
    if (ctx->p.loading == STZ_SHEAR && std::fabs(ctx->stress[ctx->lc]) < STZ_EPSILON) return true; // bmg.c:402
    return false;
}

}  // namespace

// ==================================================================================
extern "C" {

int stz_create(const stz_params *p, stz_ctx **out)
{
    if (!p || !out) return STZ_E_ARG;
    *out = nullptr;
    if (p->dim != 2 && p->dim != 3) return STZ_E_ARG;
    if (p->modes < 1 || p->modes > STZ_MAXM) return STZ_E_ARG;
    if (p->precision != STZ_FP64 && p->precision != STZ_MIXED) return STZ_E_ARG;
    if (p->loading != STZ_TENSION && p->loading != STZ_SHEAR) return STZ_E_ARG;
    int n[3] = {p->n[0], p->n[1], p->dim == 2 ? 1 : p->n[2]};
    if (!stz_solve_supported(p->dim, n, p->precision)) return STZ_E_ARG;
    stz_ctx *ctx = new (std::nothrow) stz_ctx();
    if (!ctx) return STZ_E_ARG;
    ctx->p = *p;
    ctx->p.n[2] = n[2];
    ctx->dim = p->dim;
    for (int a = 0; a < 3; ++a) ctx->n[a] = n[a];
    ctx->C = p->dim == 2 ? 3 : 6;
    ctx->S = p->dim == 2 ? 2 : 5;
    ctx->M = p->modes;
    ctx->N = (long long)n[0] * n[1] * n[2];
    ctx->lc = p->loading == STZ_TENSION ? 0 : (ctx->C == 3 ? 2 : 5);
    ctx->prec = p->precision;
    ctx->rsz = p->precision == STZ_FP64 ? 8 : 4;
    ctx->h0 = n[0] / 2 + 1;
    const int per_line = (int)(128 / (2 * ctx->rsz));   // complex values per 128-byte line
    ctx->hp = (ctx->h0 + per_line - 1) / per_line * per_line;
    ctx->nblk = (ctx->N + STZ_BLK - 1) / STZ_BLK;
    ctx->ngrp = ((long long)ctx->M * ctx->nblk + STZ_GRP - 1) / STZ_GRP;
    derive(ctx);
    ctx->rng.seed((unsigned long)p->seed);     // STZ/init.c:143-147 at rank 0

    auto fail = [&](int rc) {
        stz_destroy(ctx);
        return rc;
    };
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(STZ_E_HIP);
    const size_t N = (size_t)ctx->N, R = ctx->rsz;
    const size_t rows = (size_t)n[1] * n[2];
    int rc = STZ_OK;
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->eps0, ctx->C * N * 8);
    rc = rc ? rc : dalloc(ctx, &ctx->sigma, ctx->C * N * R);
    rc = rc ? rc : dalloc(ctx, &ctx->q0, ctx->M * N * R);
    rc = rc ? rc : dalloc(ctx, &ctx->deps, (size_t)ctx->M * ctx->S * N * R);
    rc = rc ? rc : dalloc(ctx, &ctx->deps_last, ctx->S * N * R);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->perm, N * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->temp, N * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->tlast, N * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->tloc, N * 8);
    rc = rc ? rc : dalloc(ctx, &ctx->spec, (size_t)ctx->C * rows * ctx->hp * 2 * R);
    rc = rc ? rc : dalloc(ctx, &ctx->minq, N * R);
    rc = rc ? rc : dalloc(ctx, &ctx->q0eff, ctx->M * N * R);
    rc = rc ? rc : dalloc(ctx, &ctx->scache, N * R);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->argm, N);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->bsum, (size_t)ctx->ngrp * STZ_GRP * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->gsum, ctx->ngrp * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->csum, (size_t)((ctx->ngrp + STZ_CHUNK - 1) / STZ_CHUNK) * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->ccnt, (size_t)((ctx->nblk + STZ_CHUNK - 1) / STZ_CHUNK) * sizeof(int));
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->res, 64 * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->spart, rows * ctx->C * 8);
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->acnt, ctx->nblk * sizeof(int));
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->aoff, ctx->nblk * sizeof(int));
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->alist, N * sizeof(int));
    rc = rc ? rc : dalloc(ctx, (void **)&ctx->socp, ctx->nblk * 8);
    if (rc) return fail(rc);
    if (hipHostMalloc((void **)&ctx->hres, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) return fail(STZ_E_HIP);
    std::memset(ctx->hres, 0, 64 * sizeof(double));
    // host buffers that kernels read directly (fine-grained, coherent)
    const unsigned hflags = hipHostMallocCoherent | hipHostMallocMapped;
    const size_t ms = (size_t)ctx->M * ctx->S;
    // dice, the M*S Gaussians of a glass site's redraw, the M uniforms of a crystal-like one
    if (hipHostMalloc((void **)&ctx->hspec, (1 + ms + ctx->M) * sizeof(double), hflags) != hipSuccess)
        return fail(STZ_E_HIP);
    if (hipHostMalloc((void **)&ctx->hath, (size_t)ctx->acap_max * ms * sizeof(double), hflags) != hipSuccess)
        return fail(STZ_E_HIP);
    if (hipHostGetDevicePointer((void **)&ctx->dspec_dev, ctx->hspec, 0) != hipSuccess) return fail(STZ_E_HIP);
    if (hipHostGetDevicePointer((void **)&ctx->dath_dev, ctx->hath, 0) != hipSuccess) return fail(STZ_E_HIP);
    std::memset(ctx->hspec, 0, (1 + ms + ctx->M) * sizeof(double));
    rc = ensure_dnew(ctx, (size_t)ctx->M * ctx->S);
    if (rc) return fail(rc);
    rc = stz_fft_setup(ctx);
    if (rc) return fail(rc);
    if (invalidate_cache(ctx)) return fail(STZ_E_HIP);
    // tlast = -100*tTemp (STZ/init.c:312)
    hipLaunchKernelGGL((k_fill<double>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, ctx->tlast,
                       (long long)N, -100 * ctx->k.tTemp);
    // Tlocal = T__K (STZ/init.c:470-476)
    hipLaunchKernelGGL((k_fill<double>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, ctx->tloc,
                       (long long)N, p->temperature);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(STZ_E_HIP);
    *out = ctx;
    return STZ_OK;
}

void stz_destroy(stz_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    stz_resident_free(ctx);
    stz_local_release(ctx);
    stz_inhom_free(ctx);
    void *ptrs[] = {ctx->q0eff, ctx->scache, ctx->eps0, ctx->sigma, ctx->q0, ctx->q1, ctx->deps, ctx->deps_last, ctx->perm, ctx->temp, ctx->tlast, ctx->tloc,
                    ctx->spec, ctx->minq, ctx->argm, ctx->bsum, ctx->gsum, ctx->csum, ctx->ccnt, ctx->res, ctx->spart, ctx->mro,
                    ctx->acnt, ctx->aoff, ctx->alist, ctx->dnew, ctx->socp, ctx->gneg};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    stz_fft_free(ctx);
    for (int i = 0; i < 24; ++i)
        if (ctx->tev[i]) (void)hipEventDestroy(ctx->tev[i]);
    if (ctx->hres) (void)hipHostFree(ctx->hres);
    if (ctx->hspec) (void)hipHostFree(ctx->hspec);
    if (ctx->hath) (void)hipHostFree(ctx->hath);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *stz_last_error(const stz_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"

namespace {

// modes of a crystal-like MRO family (STZ/bmg.h:30-32)
int mro_modes(int t) { return t == 1 ? 6 : (t == 2 ? 2 : 12); }

// Energy_Init (STZ/init.c:181-198), DStrain (STZ/elastic.c:60-138) and the initial solve
// (STZ/bmg.c:117-127) after the MRO map: glass voxels draw two (2-D) Gaussians per mode;
// a crystal-like voxel's mode 0 lies along its orientation, modes 1..XMODES-1 are
// rotations by 2*theta_m with (1 + 0.2(2U - 1)) factors (uniforms taken left operand
// first, the order g++ evaluates the reference's expressions in; the cos/sin pair of one
// angle from one sincos call, as g++ compiles it), later modes have zero strain.
int init_fields(stz_ctx *ctx, const std::vector<double> &theta)
{
    const long long N = ctx->N;
    std::vector<double> buf((size_t)N * ctx->S);
    for (int m = 0; m < ctx->M; ++m) {
        for (long long v = 0; v < N; ++v) buf[v] = ctx->rng.gaussian(ctx->p.q0_std) + ctx->p.q0_mean;
        int rc = put_real(ctx, ctx->q0, buf.data(), (size_t)N, (size_t)m * N);
        if (rc) return rc;
        STZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<double> first;                 // mode 0 of the crystal-like voxels: (xx, xy)
    if (ctx->nmro) first.assign((size_t)2 * N, 0.0);
    const double g0 = ctx->p.gamma0;
    for (int m = 0; m < ctx->M; ++m) {
        for (long long v = 0; v < N; ++v) {
            const int t = ctx->nmro ? ctx->mro_h[v] : 0;
            if (t == 0) {
                for (int k = 0; k < ctx->S; ++k) buf[(size_t)k * N + v] = ctx->k.half_gamma0 * ctx->rng.gaussian(1.0);
                continue;
            }
            double xx, xy;
            const int XM = mro_modes(t);
            if (m == 0) {
                const double amp = t == 2 ? 0.8 * g0 : 0.5 * g0;
                xx = amp * std::cos(theta[v]);
                xy = amp * std::sin(theta[v]);
                first[2 * v] = xx;
                first[2 * v + 1] = xy;
            } else if (m < XM) {
                const double a0 = first[2 * v], a2 = first[2 * v + 1];
                const double th = m * 2. * STZ_PI_REF / XM * (1. + 0.2 * (2 * ctx->rng.uniform() - 1.));
                double s2, c2;
                sincos(2 * th, &s2, &c2);
                const double u1 = ctx->rng.uniform();
                const double u2 = ctx->rng.uniform();
                xx = a0 * (1. + 0.2 * (2 * u1 - 1.)) * c2 + a2 * (1. + 0.2 * (2 * u2 - 1.)) * s2;
                const double u3 = ctx->rng.uniform();
                const double u4 = ctx->rng.uniform();
                xy = -a0 * (1. + 0.2 * (2 * u3 - 1.)) * s2 + a2 * (1. + 0.2 * (2 * u4 - 1.)) * c2;
            } else {
                xx = 0.;
                xy = 0.;
            }
            buf[v] = xx;
            buf[(size_t)N + v] = xy;
        }
        int rc = put_real(ctx, ctx->deps, buf.data(), (size_t)N * ctx->S, (size_t)m * ctx->S * N);
        if (rc) return rc;
        STZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (ctx->q1)   // Q1 = Q0 (STZ/init.c:189)
        STZ_HIP(hipMemcpyAsync(ctx->q1, ctx->q0, (size_t)ctx->M * N * ctx->rsz, hipMemcpyDeviceToDevice, ctx->stream));
    const int rc = invalidate_cache(ctx);   // new Q0 and strains
    if (rc) return rc;
    for (int c = 0; c < 6; ++c) ctx->epsavg[c] = 0.0;
    return solve(ctx);   // STZ/bmg.c:117-127
}

}  // namespace

// installs ctx->mro_h on the device (no crystal-like voxel: none, all glass)
int stz_set_mro_map(stz_ctx *ctx)
{
    long long n = 0;
    for (uint8_t t : ctx->mro_h) n += t != 0;
    if (n && ctx->dim != 2) {
        ctx->err = "crystal-like MRO families are defined for 2-D grids only";
        return STZ_E_UNSUPPORTED;
    }
    if (ctx->mro) STZ_HIP(hipFree(ctx->mro));
    ctx->mro = nullptr;
    ctx->nmro = n;
    if (!n) {
        ctx->mro_h.clear();
        if (ctx->q1) STZ_HIP(hipFree(ctx->q1));
        ctx->q1 = nullptr;
        return STZ_OK;
    }
    // only crystal-like transforms change Q0 (STZ/elastic.c:446-506), so only MRO contexts
    // keep the initial barriers Q1; the caller fills it
    if (!ctx->q1) {
        const int rc = dalloc(ctx, &ctx->q1, (size_t)ctx->M * ctx->N * ctx->rsz);
        if (rc) return rc;
    }
    STZ_HIP(hipMalloc((void **)&ctx->mro, (size_t)ctx->N));
    STZ_HIP(hipMemcpy(ctx->mro, ctx->mro_h.data(), (size_t)ctx->N, hipMemcpyHostToDevice));
    ctx->acap = 0;              // avalanches: the host draws per voxel by type
    return STZ_OK;
}

extern "C" {

int stz_init_state(stz_ctx *ctx)
{
    if (!ctx) return STZ_E_ARG;
    stz_rng_sync(ctx);
    const long long N = ctx->N;
    // MRO map, InitMRO 0 (STZ/init.c:350-397): one uniform per voxel; six-fold
    // crystal-like MRO when it is <= xMROFrac, oriented by a second uniform
    ctx->mro_h.assign((size_t)N, 0);
    std::vector<double> theta((size_t)N, 0.0);
    for (long long v = 0; v < N; ++v) {
        const double u = ctx->rng.uniform();
        if (u <= ctx->p.mro_fraction) {
            ctx->mro_h[v] = 1;
            theta[v] = ctx->rng.uniform() * STZ_PI_REF - STZ_PI_REF / 2.0;
        }
    }
    int rc = stz_set_mro_map(ctx);
    if (rc) return rc;
    return init_fields(ctx, theta);
}

int stz_init_state_mro(stz_ctx *ctx, const int *type, const double *theta)
{
    if (!ctx || !type || !theta) return STZ_E_ARG;
    stz_rng_sync(ctx);
    const long long N = ctx->N;
    ctx->mro_h.assign((size_t)N, 0);
    std::vector<double> th(theta, theta + N);
    for (long long v = 0; v < N; ++v) {
        if (type[v] < 0 || type[v] > 3) {
            ctx->err = "MRO type outside {0, 1, 2, 3}";      // STZ/elastic.c:131-136
            return STZ_E_ARG;
        }
        ctx->mro_h[v] = (uint8_t)type[v];
    }
    // InitMRO 1 (STZ/init.c:420-428): ten-fold voxels become two-fold with probability 0.5
    for (long long v = 0; v < N; ++v)
        if (ctx->mro_h[v] == 3 && ctx->rng.uniform() < 0.5) ctx->mro_h[v] = 2;
    int rc = stz_set_mro_map(ctx);
    if (rc) return rc;
    return init_fields(ctx, th);
}

int stz_init_state_device(stz_ctx *ctx, unsigned long long seed, const double *spheres, int nsph, double inc_mean,
                          double inc_std, long long *inside)
{
    if (!ctx || nsph < 0 || (nsph > 0 && !spheres)) return STZ_E_ARG;
    int rc = stz_init_device_launch(ctx, seed, spheres, nsph, inc_mean, inc_std, inside);
    if (rc) return rc;
    rc = invalidate_cache(ctx);
    if (rc) return rc;
    for (int c = 0; c < 6; ++c) ctx->epsavg[c] = 0.0;
    return solve(ctx);   // STZ/bmg.c:117-127
}

int stz_upload(stz_ctx *ctx, int field, const double *host)
{
    if (!ctx || !host) return STZ_E_ARG;
    const size_t n = field_count(ctx, field);
    if (n == 0 || field == STZ_F_MINQ) return STZ_E_ARG;
    int rc = STZ_OK;
    switch (field) {
    case STZ_F_EPS0: STZ_HIP(hipMemcpyAsync(ctx->eps0, host, n * 8, hipMemcpyHostToDevice, ctx->stream)); break;
    case STZ_F_SIGMA:
        ctx->pend_comp = -1;
        ctx->pend_delta = 0.0;
        for (int c = 0; c < 6; ++c) ctx->soff[c] = 0.0;
        rc = put_real(ctx, ctx->sigma, host, n);
        break;
    case STZ_F_Q0:
        rc = put_real(ctx, ctx->q0, host, n);
        if (!rc && ctx->q1) rc = put_real(ctx, ctx->q1, host, n);   // the barriers a print step restores
        break;
    case STZ_F_DEPS: rc = put_real(ctx, ctx->deps, host, n); break;
    case STZ_F_DEPS_LAST: rc = put_real(ctx, ctx->deps_last, host, n); break;
    case STZ_F_PERM: STZ_HIP(hipMemcpyAsync(ctx->perm, host, n * 8, hipMemcpyHostToDevice, ctx->stream)); break;
    case STZ_F_TEMP: STZ_HIP(hipMemcpyAsync(ctx->temp, host, n * 8, hipMemcpyHostToDevice, ctx->stream)); break;
    case STZ_F_TLAST: STZ_HIP(hipMemcpyAsync(ctx->tlast, host, n * 8, hipMemcpyHostToDevice, ctx->stream)); break;
    case STZ_F_TLOCAL: STZ_HIP(hipMemcpyAsync(ctx->tloc, host, n * 8, hipMemcpyHostToDevice, ctx->stream)); break;
    default: return STZ_E_ARG;
    }
    if (rc) return rc;
    if (field == STZ_F_EPS0 || field == STZ_F_SIGMA) ctx->nlocal = 1ll << 60;   // full solve before a local update
    if (field != STZ_F_EPS0 && field != STZ_F_SIGMA && field != STZ_F_TLOCAL) {
        rc = invalidate_cache(ctx);
        if (rc) return rc;
    }
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    return STZ_OK;
}

int stz_download(stz_ctx *ctx, int field, double *host)
{
    if (!ctx || !host) return STZ_E_ARG;
    const size_t n = field_count(ctx, field);
    if (n == 0) return STZ_E_ARG;
    switch (field) {
    case STZ_F_EPS0: STZ_HIP(hipMemcpyAsync(host, ctx->eps0, n * 8, hipMemcpyDeviceToHost, ctx->stream)); break;
    case STZ_F_SIGMA: {
        int rc = get_real(ctx, ctx->sigma, host, n);
        if (rc) return rc;
        const size_t N = (size_t)ctx->N;
        if (ctx->prec == STZ_FP64) {
            if (ctx->pend_comp >= 0)
                for (size_t v = 0; v < N; ++v) host[(size_t)ctx->pend_comp * N + v] += ctx->pend_delta;
        } else {
            for (int c = 0; c < ctx->C; ++c)
                for (size_t v = 0; v < N; ++v) host[(size_t)c * N + v] += ctx->soff[c];
        }
        return STZ_OK;
    }
    case STZ_F_Q0: return get_real(ctx, ctx->q0, host, n);
    case STZ_F_DEPS: return get_real(ctx, ctx->deps, host, n);
    case STZ_F_DEPS_LAST: return get_real(ctx, ctx->deps_last, host, n);
    case STZ_F_MINQ: return get_real(ctx, ctx->minq, host, n);
    case STZ_F_PERM: STZ_HIP(hipMemcpyAsync(host, ctx->perm, n * 8, hipMemcpyDeviceToHost, ctx->stream)); break;
    case STZ_F_TEMP: STZ_HIP(hipMemcpyAsync(host, ctx->temp, n * 8, hipMemcpyDeviceToHost, ctx->stream)); break;
    case STZ_F_TLAST: STZ_HIP(hipMemcpyAsync(host, ctx->tlast, n * 8, hipMemcpyDeviceToHost, ctx->stream)); break;
    case STZ_F_TLOCAL: STZ_HIP(hipMemcpyAsync(host, ctx->tloc, n * 8, hipMemcpyDeviceToHost, ctx->stream)); break;
    case STZ_F_MRO:
        for (size_t v = 0; v < n; ++v) host[v] = ctx->nmro ? (double)ctx->mro_h[v] : 0.0;
        return STZ_OK;
    default: return STZ_E_ARG;
    }
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    return STZ_OK;
}

int stz_set_scalars(stz_ctx *ctx, const double epsavg[6], const double stress[6], double totime)
{
    if (!ctx) return STZ_E_ARG;
    for (int c = 0; c < 6; ++c) {
        if (epsavg) ctx->epsavg[c] = epsavg[c];
        if (stress) ctx->stress[c] = stress[c];
    }
    if (totime != ctx->totime) {
        ctx->totime = totime;
        const int rc = invalidate_cache(ctx);   // settled softening entries assume a clock that only advances
        if (rc) return rc;
    }
    return STZ_OK;
}

int stz_get_scalars(stz_ctx *ctx, double epsavg[6], double stress[6], double *totime)
{
    if (!ctx) return STZ_E_ARG;
    for (int c = 0; c < 6; ++c) {
        if (epsavg) epsavg[c] = ctx->epsavg[c];
        if (stress) stress[c] = ctx->stress[c];
    }
    if (totime) *totime = ctx->totime;
    return STZ_OK;
}

int stz_stress_solve(stz_ctx *ctx) { return ctx ? solve(ctx) : STZ_E_ARG; }

}  // extern "C"

int stz_invalidate_derived(stz_ctx *ctx)
{
    const int rc = invalidate_cache(ctx);
    if (rc) return rc;
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    return STZ_OK;
}

extern "C" {

int stz_stress_update_local(stz_ctx *ctx, int n, const int *sites)
{
    if (!ctx || n < 0 || n > STZ_LOCAL_MAXSRC || (n > 0 && !sites)) return STZ_E_ARG;
    for (int i = 0; i < n; ++i) {
        if (sites[i] < 0 || sites[i] >= ctx->N) return STZ_E_ARG;
        for (int j = 0; j < i; ++j)
            if (sites[j] == sites[i]) return STZ_E_ARG;     // each site's recorded strain is added once
    }
    if (ctx->lamx) {
        ctx->err = "the event-local stress update assumes uniform moduli";
        return STZ_E_UNSUPPORTED;
    }
    if (ctx->nlocal >= (1ll << 59)) {
        ctx->err = "the stress field has not been solved since its inputs were replaced";
        return STZ_E_ARG;
    }
    int rc = stz_local_prepare(ctx);
    if (rc) return rc;
    if (ctx->prec == STZ_FP64 && ctx->pend_comp >= 0) {
        rc = flush_pending(ctx);              // the increment is counted in usum
        if (rc) return rc;
    }
    double uni[6];
    stz_local_uniform(ctx, uni);
    STZ_HIP(hipMemsetAsync(ctx->res + 15, 0, sizeof(double), ctx->stream));
    if (n > 0) STZ_HIP(hipMemcpyAsync(ctx->lsrc, sites, (size_t)n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    rc = stz_local_launch(ctx, n, uni);
    if (rc) return rc;
    STZ_HIP(hipMemcpyAsync(ctx->hres, ctx->res, 16 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    int nan;
    std::memcpy(&nan, ctx->hres + 15, sizeof(int));
    if (nan) {
        ctx->err = "NaN in the stress update";
        return STZ_E_NAN;
    }
    for (int c = 0; c < ctx->C; ++c) ctx->stress[c] = ctx->hres[c];
    for (int c = 0; c < 6; ++c) {
        ctx->soff[c] = 0.0;                   // the update rewrote the stress field
        ctx->ebase[c] = ctx->epsavg[c];
        ctx->usum[c] = 0.0;
    }
    ctx->pend_comp = -1;
    ctx->pend_delta = 0.0;
    ctx->nlocal++;
    ctx->nsolve++;
    return STZ_OK;
}

int stz_displacement(stz_ctx *ctx, double *host_u)
{
    if (!ctx || !host_u) return STZ_E_ARG;
    if (ctx->lamx) {
        // the displacement comes from the eigenstrain the fixed point ends on: solve first,
        // as the reference's print-step InhomogeneousStressSolver(uflag=2) call does
        const int rc = solve(ctx);
        if (rc) return rc;
    }
    const size_t bytes = (size_t)ctx->dim * (size_t)ctx->N * 8;
    double *u = nullptr;
    STZ_HIP(hipMalloc((void **)&u, bytes));
    int rc = stz_displacement_launch(ctx, u);
    if (!rc && hipMemcpyAsync(host_u, u, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) rc = STZ_E_HIP;
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = STZ_E_HIP;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(u);
    if (rc == STZ_E_HIP && ctx->err.empty()) ctx->err = "displacement: HIP error";
    return rc;
}

int stz_set_heat_conduct(stz_ctx *ctx, int on)
{
    if (!ctx) return STZ_E_ARG;
    ctx->heat = on != 0;
    ctx->tune_epoch++;
    return STZ_OK;
}

int stz_heat_conduct(stz_ctx *ctx, double dt)
{
    if (!ctx || !(dt >= 0)) return STZ_E_ARG;
    return heat_step(ctx, dt);
}

int stz_set_solver(stz_ctx *ctx, int mode, long long resync, int local_max)
{
    if (!ctx || (mode != STZ_SOLVER_FFT && mode != STZ_SOLVER_LOCAL) || resync < 1 || local_max < 1 ||
        local_max > STZ_LOCAL_MAXSRC)
        return STZ_E_ARG;
    if (mode == STZ_SOLVER_LOCAL && ctx->lamx) {
        ctx->err = "the event-local stress update assumes uniform moduli";
        return STZ_E_UNSUPPORTED;
    }
    if (mode == STZ_SOLVER_LOCAL) {
        const int rc = stz_local_prepare(ctx);
        if (rc) return rc;
    }
    ctx->solver_mode = mode;
    ctx->local_resync = resync;
    ctx->local_max = local_max;
    ctx->tune_epoch++;                        // captured step graphs hold the launch sequence
    return STZ_OK;
}

int stz_rates(stz_ctx *ctx, long long *n_negative, double *total_rate)
{
    if (!ctx) return STZ_E_ARG;
    long long nn;
    double R;
    int rc = rates(ctx, &nn, &R);
    if (rc) return rc;
    if (n_negative) *n_negative = nn;
    if (total_rate) *total_rate = R;
    return STZ_OK;
}

int stz_select(stz_ctx *ctx, double dice, int *site, int *mode, double *q)
{
    if (!ctx || !site || !mode) return STZ_E_ARG;
    double qq;
    int rc = select_event(ctx, dice, site, mode, &qq);
    if (rc) return rc;
    if (q) *q = qq;
    return STZ_OK;
}

int stz_apply_event(stz_ctx *ctx, int site, int mode, const double *deps_new)
{
    if (!ctx || !deps_new || site < 0 || site >= ctx->N || mode < 0 || mode >= ctx->M) return STZ_E_ARG;
    int rc = apply_event(ctx, site, mode, deps_new);
    if (rc) return rc;
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    return STZ_OK;
}

int stz_athermal_scan(stz_ctx *ctx, int *n)
{
    if (!ctx || !n) return STZ_E_ARG;
    return athermal_scan(ctx, n);
}

int stz_athermal_list(stz_ctx *ctx, int *sites, int *modes, int cap)
{
    if (!ctx || !sites) return STZ_E_ARG;
    const int n = ctx->n_athermal < cap ? ctx->n_athermal : cap;
    if (n <= 0) return STZ_OK;
    STZ_HIP(hipMemcpyAsync(sites, ctx->alist, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    if (modes) {
        std::vector<uint8_t> am((size_t)ctx->N);
        STZ_HIP(hipMemcpyAsync(am.data(), ctx->argm, (size_t)ctx->N, hipMemcpyDeviceToHost, ctx->stream));
        STZ_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; ++i) modes[i] = am[sites[i]];
    }
    return STZ_OK;
}

int stz_athermal_apply(stz_ctx *ctx, const double *deps_new, double *soc_elast)
{
    if (!ctx || (!deps_new && ctx->n_athermal > 0)) return STZ_E_ARG;
    return athermal_apply(ctx, deps_new, soc_elast);
}

int stz_add_uniform_stress(stz_ctx *ctx, int comp, double delta)
{
    if (!ctx || comp < 0 || comp >= ctx->C) return STZ_E_ARG;
    return add_uniform(ctx, comp, delta);
}

int stz_elastic_energy(stz_ctx *ctx, double *e)
{
    if (!ctx || !e) return STZ_E_ARG;
    int rc = upload_soff(ctx);
    if (rc) return rc;
    rc = stz_energy_launch(ctx);
    if (rc) return rc;
    std::vector<double> part((size_t)ctx->nblk);
    STZ_HIP(hipMemcpyAsync(part.data(), ctx->socp, ctx->nblk * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (long long i = 0; i < ctx->nblk; ++i) s += part[(size_t)i];
    *e = s;
    return STZ_OK;
}

int stz_set_print(stz_ctx *ctx, int printperiod, double endstrain)
{
    if (!ctx) return STZ_E_ARG;
    ctx->print_on = 1;
    ctx->printperiod = printperiod;
    ctx->endstrain = endstrain;
    return STZ_OK;
}

int stz_step_kmc(stz_ctx *ctx, stz_step_result *res)
{
    if (!ctx || !res) return STZ_E_ARG;
    stz_rng_sync(ctx);
    std::memset(res, 0, sizeof(*res));
    res->site = res->mode = -1;
    ctx->step++;
    const size_t ms = (size_t)ctx->M * ctx->S;
    // The random numbers a thermal event would consume -- the dice (STZ/evolution.c:142)
    // and the M*S redraw values (STZ/elastic.c:508-512) -- and those of an athermal
    // avalanche of up to acap voxels (row-major, M*S each) are drawn before the step
    // kind is known; afterwards the stream is set to exactly what the reference consumed.
    const StzRng s0 = ctx->rng;
    ctx->hspec[0] = ctx->rng.uniform();
    const StzRng s_dice = ctx->rng;
    draw_voxel_deps(ctx, ctx->hspec + 1);
    const StzRng s_th = ctx->rng;
    StzRng s_thx = s_th;
    if (ctx->nmro) {
        // a crystal-like site redraws its M barriers from M uniforms instead (STZ/elastic.c:446-506)
        ctx->rng = s_dice;
        for (int m = 0; m < ctx->M; ++m) ctx->hspec[1 + ms + m] = ctx->rng.uniform();
        s_thx = ctx->rng;
    }
    int rc = stz_rates_launch(ctx, 1);            // rates + group sums + step kind on the device
    if (rc) return rc;
    ctx->pend_comp = -1;
    ctx->pend_delta = 0.0;
    ctx->rng = s0;                                // host drawing overlaps the rate pass
    for (long long i = 0; i < ctx->acap; ++i) draw_voxel_deps(ctx, ctx->hath + (size_t)i * ms);
    ctx->rng = s0;
    rc = stz_select_launch(ctx, 0.0, 1);          // thermal: selection + transform
    if (rc) return rc;
    rc = stz_athermal_fused_launch(ctx, 1);       // athermal: compaction + transforms
    if (rc) return rc;
    if (!ctx->lamx) {
        rc = stz_solve_launch(ctx, 1);            // thermal or athermal: stress solve
        if (rc) return rc;
    }
    STZ_HIP(hipMemcpyAsync(ctx->hres, ctx->res, 32 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    rc = spin_sync(ctx);
    if (rc) return rc;
    const double R = ctx->hres[8];
    const long long nneg = (long long)ctx->hres[9];
    int kind;
    std::memcpy(&kind, ctx->hres + RES_KIND, sizeof(int));
    ctx->rate_total = R;
    res->n_negative = nneg;
    res->total_rate = R;
    acc_rates(ctx);
    if (kind == KIND_FAIL) {
        ctx->err = "KMC failed to locate an activation mode";
        return STZ_E_SELECT;
    }
    if ((kind == KIND_THERMAL || kind == KIND_ATHERMAL) && ctx->lamx) {
        rc = solve(ctx);                          // the inhomogeneous loop decides on the host
        if (rc) return rc;
    } else if (kind == KIND_THERMAL || kind == KIND_ATHERMAL) {
        int nan;
        std::memcpy(&nan, ctx->hres + 15, sizeof(int));
        if (nan) {
            ctx->err = "NaN in the stress solve";
            return STZ_E_NAN;
        }
        acc_solve(ctx);
        ctx->nsolve++;
        for (int c = 0; c < ctx->C; ++c) ctx->stress[c] = ctx->hres[c];
        for (int c = 0; c < 6; ++c) ctx->soff[c] = 0.0;   // the solve rewrote the stress field
        stz_local_mark_solved(ctx);
    }
    if (kind == KIND_THERMAL) {
        acc_select(ctx);
        const int tsite = (int)ctx->hres[10];
        ctx->rng = (ctx->nmro && ctx->mro_h[tsite]) ? s_thx : s_th;
        ctx->avlan = 0;                           // UpdateSOC (STZ/evolution.c:5-45)
        ctx->soc_elast = 0.0;
        const double t = 1 / R;                   // STZ/evolution.c:134
        res->dt_kmc = t;
        ctx->kstep++;
        ctx->ntrans++;
        res->kind = STZ_STEP_THERMAL;
        res->site = (int)ctx->hres[10];
        res->mode = (int)ctx->hres[11];
        res->q_chosen = ctx->hres[12];
        rc = load_thermal(ctx, t);
        if (rc) return rc;
        ctx->totime += t;                         // STZ/evolution.c:317
        if (ctx->heat) {
            rc = heat_step(ctx, STZ_THERMAL_DIFFUSIVITY * t);   // STZ/evolution.c:247-249
            if (rc) return rc;
        }
    } else if (kind == KIND_ATHERMAL || kind == KIND_ATH_OVERFLOW) {
        // athermal_plasticity (STZ/evolution.c:47-100)
        ctx->astep++;
        const long long n = (long long)ctx->hres[13];
        double soc = ctx->hres[14];
        ctx->rng = s0;
        if (kind == KIND_ATHERMAL) {
            for (long long i = 0; i < n; ++i) draw_voxel_deps(ctx, ctx->hath);   // advance the stream
        } else {
            // more voxels than pre-drawn: transform the rest with the continuing stream
            ctx->hbuf.resize((size_t)n * ms);
            if (ctx->nmro) {
                // glass voxels draw M*S Gaussians, crystal-like ones M uniforms, in list order
                std::vector<int> list((size_t)n);
                STZ_HIP(hipMemcpyAsync(list.data(), ctx->alist, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                                       ctx->stream));
                STZ_HIP(hipStreamSynchronize(ctx->stream));
                for (long long i = 0; i < n; ++i) {
                    double *slot = ctx->hbuf.data() + (size_t)i * ms;
                    if (ctx->mro_h[list[i]])
                        for (int m = 0; m < ctx->M; ++m) slot[m] = ctx->rng.uniform();
                    else
                        draw_voxel_deps(ctx, slot);
                }
            } else {
                for (long long i = 0; i < n; ++i) draw_voxel_deps(ctx, ctx->hbuf.data() + (size_t)i * ms);
            }
            const long long rest = n - ctx->acap;
            rc = ensure_dnew(ctx, (size_t)rest * ms);
            if (rc) return rc;
            STZ_HIP(hipMemcpyAsync(ctx->dnew, ctx->hbuf.data() + (size_t)ctx->acap * ms, (size_t)rest * ms * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
            rc = upload_soff(ctx);
            if (rc) return rc;
            rc = stz_athermal_apply_range(ctx, ctx->acap, rest);
            if (rc) return rc;
            std::vector<double> part((size_t)((rest + 255) / 256));
            STZ_HIP(hipMemcpyAsync(part.data(), ctx->socp, part.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            rc = solve(ctx);
            if (rc) return rc;
            for (double x : part) soc += x;
        }
        ctx->ntrans += n;
        ctx->avlan += n;
        ctx->soc_elast += soc;
        athermal_correction(ctx);
        // pre-draw enough redraws for the avalanches seen recently
        long long want = 2 * n;
        if (want < 8) want = 8;
        if (want > ctx->acap_max) want = ctx->acap_max;
        if (want > ctx->acap && !ctx->nmro) ctx->acap = want;
        res->kind = STZ_STEP_ATHERMAL;
        res->n_transformed = (int)n;
        res->soc_elast = soc;
    } else {
        ctx->rng = s0;                            // nothing consumed
        ctx->avlan = 0;                           // UpdateSOC
        ctx->soc_elast = 0.0;
        res->dt_kmc = 1 / R;
        ctx->estep++;
        ctx->totime += ctx->k.dt;                 // STZ/evolution.c:324
        if (ctx->heat) {
            rc = heat_step(ctx, STZ_THERMAL_DIFFUSIVITY * ctx->k.dt);   // STZ/evolution.c:325-327
            if (rc) return rc;
        }
        rc = load_elastic(ctx);
        if (rc) return rc;
        res->kind = STZ_STEP_ELASTIC;
    }
    int term = 0;
    if (ctx->print_on) {
        // STZ/bmg.c:181-187; a period <= 0 makes the test true at every step, as there
        const double cur = ctx->epsavg[ctx->lc];
        if (std::fabs(cur - ctx->strainhist_last) >= (ctx->printperiod * ctx->p.strainincrement * 0.99)) {
            ctx->strainhist_last = cur;           // strainhist[N_print-1], STZ/bmg.c:193-201
            res->printed = 1;
        } else {
            term = terminated(ctx) ? 1 : 0;
        }
    }
    res->totime = ctx->totime;
    for (int c = 0; c < 6; ++c) {
        res->epsavg[c] = ctx->epsavg[c];
        res->stress[c] = ctx->stress[c];
    }
    return term;
}

int stz_step_print(stz_ctx *ctx, stz_step_result *res)
{
    if (!ctx || !res) return STZ_E_ARG;
    int rc;
    if (ctx->q1) {
        // Q0 <- Q1 (STZ/bmg.c:215-223): the barriers crystal-like transforms redrew return to
        // their initial draws (glass barriers never change, so only MRO contexts keep Q1)
        STZ_HIP(hipMemcpyAsync(ctx->q0, ctx->q1, (size_t)ctx->M * ctx->N * ctx->rsz, hipMemcpyDeviceToDevice,
                               ctx->stream));
        rc = invalidate_cache(ctx);
        if (rc) return rc;
    }
    rc = solve(ctx);                              // print-step re-solve, STZ/bmg.c:289-295
    if (rc) return rc;
    for (int c = 0; c < 6; ++c) res->stress[c] = ctx->stress[c];
    return terminated(ctx) ? 1 : 0;               // STZ/bmg.c:382-410
}

int stz_step(stz_ctx *ctx, stz_step_result *res)
{
    int rc = stz_step_kmc(ctx, res);
    if (rc < 0 || !res->printed) return rc;
    return stz_step_print(ctx, res);
}

int stz_rng_seed(stz_ctx *ctx, unsigned long seed)
{
    if (!ctx) return STZ_E_ARG;
    ctx->rng.seed(seed);
    stz_rng_replaced(ctx);
    return STZ_OK;
}

int stz_rng_uniform(stz_ctx *ctx, double *u)
{
    if (!ctx || !u) return STZ_E_ARG;
    stz_rng_sync(ctx);
    *u = ctx->rng.uniform();
    return STZ_OK;
}

int stz_rng_gaussian(stz_ctx *ctx, double sigma, double *g)
{
    if (!ctx || !g) return STZ_E_ARG;
    stz_rng_sync(ctx);
    *g = ctx->rng.gaussian(sigma);
    return STZ_OK;
}

int stz_rng_get_state(stz_ctx *ctx, unsigned int words[625])
{
    if (!ctx || !words) return STZ_E_ARG;
    stz_rng_sync(ctx);
    for (int i = 0; i < 624; ++i) words[i] = ctx->rng.state[i];
    words[624] = (unsigned int)ctx->rng.pos;
    return STZ_OK;
}

int stz_rng_set_state(stz_ctx *ctx, const unsigned int words[625])
{
    if (!ctx || !words || words[624] > 624) return STZ_E_ARG;
    for (int i = 0; i < 624; ++i) ctx->rng.state[i] = words[i];
    ctx->rng.pos = (int)words[624];
    ctx->rng.count = 0;
    stz_rng_replaced(ctx);
    return STZ_OK;
}

int stz_get_print_state(stz_ctx *ctx, double out[3])
{
    if (!ctx || !out) return STZ_E_ARG;
    out[0] = ctx->strainhist_last;
    out[1] = ctx->printperiod;
    out[2] = ctx->endstrain;
    return STZ_OK;
}

int stz_set_print_state(stz_ctx *ctx, double strainhist_last)
{
    if (!ctx) return STZ_E_ARG;
    ctx->strainhist_last = strainhist_last;
    return STZ_OK;
}

int stz_timing_enable(stz_ctx *ctx, int on)
{
    if (!ctx) return STZ_E_ARG;
    if (on && !ctx->tev[0]) {
        for (int i = 0; i < 24; ++i) STZ_HIP(hipEventCreate(&ctx->tev[i]));
    }
    ctx->timing = on ? 1 : 0;
    for (int i = 0; i < 8; ++i) {
        ctx->tms[i] = ctx->sms[i] = 0.0;
        ctx->tn[i] = ctx->sn[i] = 0;
    }
    return STZ_OK;
}

int stz_timing_get(stz_ctx *ctx, double ms[8], long long n[8])
{
    if (!ctx || !ms || !n) return STZ_E_ARG;
    for (int i = 0; i < 8; ++i) {
        ms[i] = ctx->tms[i];
        n[i] = ctx->tn[i];
    }
    return STZ_OK;
}

int stz_timing_solve_get(stz_ctx *ctx, double ms[8], long long n[8])
{
    if (!ctx || !ms || !n) return STZ_E_ARG;
    for (int i = 0; i < 8; ++i) {
        ms[i] = ctx->sms[i];
        n[i] = ctx->sn[i];
    }
    return STZ_OK;
}

// Autotuner of the stress-solve tile schedule: every compiled variant of each of the
// four kernel slots is timed with HIP events on this context's own fields (the stress
// field is saved and restored), and the fastest is kept.  Results are cached per grid
// shape and precision for the process.
int stz_autotune(stz_ctx *ctx, int reps, int flags, double best_ms[4])
{
    if (!ctx || reps < 1) return STZ_E_ARG;
    static std::map<std::tuple<int, int, int, int, int>, std::array<int, 4>> cache;
    static std::mutex mu;
    const auto key = std::make_tuple(ctx->dim, ctx->n[0], ctx->n[1], ctx->n[2], ctx->prec);
    if (!(flags & 1)) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            for (int s = 0; s < 4; ++s) ctx->tune[s] = it->second[(size_t)s];
            ctx->tune_epoch++;
            if (best_ms)
                for (int s = 0; s < 4; ++s) best_ms[s] = -1.0;
            return STZ_OK;
        }
    }
    int cand[4][STZ_TUNE_MAXCAND], nc[4], maxc = 0;
    for (int s = 0; s < 4; ++s) {
        nc[s] = tune_candidates(ctx, s, cand[s]);
        maxc = std::max(maxc, nc[s]);
    }
    if (!ctx->tev[0])
        for (int i = 0; i < 24; ++i) STZ_HIP(hipEventCreate(&ctx->tev[i]));
    const size_t sbytes = (size_t)ctx->C * (size_t)ctx->N * ctx->rsz;
    void *save = nullptr;
    STZ_HIP(hipMalloc(&save, sbytes));
    STZ_HIP(hipMemcpyAsync(save, ctx->sigma, sbytes, hipMemcpyDeviceToDevice, ctx->stream));
    const int saved_timing = ctx->timing;
    const int saved_tune[4] = {ctx->tune[0], ctx->tune[1], ctx->tune[2], ctx->tune[3]};
    double t[4][STZ_TUNE_MAXCAND];
    int rc = STZ_OK;
    for (int i = 0; i < maxc && rc == STZ_OK; ++i) {
        for (int s = 0; s < 4; ++s) ctx->tune[s] = nc[s] ? cand[s][std::min(i, nc[s] - 1)] : 0;
        ctx->timing = 1;
        rc = stz_solve_launch(ctx, 0);
        if (rc || hipStreamSynchronize(ctx->stream) != hipSuccess) break;
        for (int j = 0; j < 8; ++j) {
            ctx->sms[j] = 0.0;
            ctx->sn[j] = 0;
        }
        for (int r = 0; r < reps && rc == STZ_OK; ++r) {
            rc = stz_solve_launch(ctx, 0);
            if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = STZ_E_HIP;
            acc_solve(ctx);
        }
        const double n = (double)std::max<long long>(ctx->sn[0], 1);
        t[0][i] = ctx->sms[0] / n;
        t[1][i] = (ctx->sms[1] + ctx->sms[3]) / n;
        t[2][i] = ctx->sms[2] / n;
        t[3][i] = ctx->sms[4] / n;
    }
    for (int s = 0; s < 4; ++s) {
        int best = 0;
        for (int i = 1; i < nc[s]; ++i)
            if (t[s][i] < t[s][best]) best = i;
        ctx->tune[s] = nc[s] ? cand[s][best] : 0;
        if (best_ms) best_ms[s] = nc[s] ? t[s][best] : 0.0;
    }
    if (rc) for (int s = 0; s < 4; ++s) ctx->tune[s] = saved_tune[s];
    ctx->tune_epoch++;
    (void)hipMemcpyAsync(ctx->sigma, save, sbytes, hipMemcpyDeviceToDevice, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(save);
    ctx->timing = saved_timing;
    for (int j = 0; j < 8; ++j) {
        ctx->tms[j] = ctx->sms[j] = 0.0;
        ctx->tn[j] = ctx->sn[j] = 0;
    }
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = {ctx->tune[0], ctx->tune[1], ctx->tune[2], ctx->tune[3]};
    return STZ_OK;
}

int stz_tune_candidates(stz_ctx *ctx, int slot, int out[STZ_TUNE_MAXCAND])
{
    if (!ctx || !out || slot < 0 || slot > 3) return STZ_E_ARG;
    return tune_candidates(ctx, slot, out);
}

int stz_get_tuning(stz_ctx *ctx, int out[4])
{
    if (!ctx || !out) return STZ_E_ARG;
    for (int s = 0; s < 4; ++s) out[s] = ctx->tune[s];
    return STZ_OK;
}

int stz_set_tuning(stz_ctx *ctx, const int in[4])
{
    if (!ctx || !in) return STZ_E_ARG;
    for (int s = 0; s < 4; ++s) ctx->tune[s] = in[s] < 0 ? 0 : in[s];
    ctx->tune_epoch++;
    return STZ_OK;
}

int stz_counters(stz_ctx *ctx, long long out[6])
{
    if (!ctx || !out) return STZ_E_ARG;
    out[0] = ctx->step;
    out[1] = ctx->astep;
    out[2] = ctx->kstep;
    out[3] = ctx->estep;
    out[4] = ctx->ntrans;
    out[5] = ctx->nsolve;
    return STZ_OK;
}

int stz_copy_probe(long long bytes, int reps, double *avg_ms, double *gbs)
{
    if (bytes < 16 * 256 || reps < 1 || !avg_ms || !gbs) return STZ_E_ARG;
    const long long n = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = STZ_E_HIP;
    const unsigned grid = (unsigned)((n + 255) / 256);
    float ms = 0.f;
    if (hipMalloc(&src, n * 16) != hipSuccess || hipMalloc(&dst, n * 16) != hipSuccess) goto out;
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        goto out;
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s, (double *)src, 2 * n,
                       1.0);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n);   // warm
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto out;
    *avg_ms = ms / reps;
    *gbs = 2.0 * (double)(n * 16) / (*avg_ms * 1e-3) / 1e9;
    rc = STZ_OK;
out:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

}  // extern "C"
