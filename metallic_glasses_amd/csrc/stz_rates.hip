// stz_rates.hip -- activation rates, residence-time selection, athermal scan and the
// eigenstrain update.
//
// k_rates   one pass over all voxels: SofteningFactor (STZ/init.c:711-739), ActEnergy
//           (STZ/init.c:561-690), the Q<0 count (STZ/bmg.c:147-153), the KMC rates
//           (STZ/evolution.c:120-128) -- never materialised -- reduced to FP64 sums per
//           (mode, block of 256 voxels) in a fixed tree, plus the per-voxel minimum Q
//           and its first argmin for the athermal scan (STZ/evolution.c:55-64).
// k_group   fixed-tree sums of 256 block sums.
// k_total   fixed-order running scan over the group sums -> R.
// k_select  three-level search for the first event index whose cumulative rate reaches
//           dice*R (STZ/evolution.c:150-224): groups -> blocks -> voxels, each level a
//           wavefront-shuffle scan whose last element reproduces the sums above bit for
//           bit, the voxel level re-evaluating the 256 rates of one block with the same
//           code as k_rates.
// k_apply   VoxelShearTransform (STZ/elastic.c:361-520) for one or many voxels.
#include "stz_kmc_dev.h"


namespace {

template <typename R, int DIM, bool XTAL>
__global__ __launch_bounds__(NT) void k_rates(RateArgs<R> a)
{
    rates_block<R, DIM, XTAL>(a, blockIdx.x);
}

// Fixed-tree sums of STZ_GRP consecutive block sums, STZ_GRP_PER_BLOCK groups per workgroup
__global__ __launch_bounds__(NT) void k_group(const double *__restrict__ bsum, long long nb, double *__restrict__ gsum,
                                              long long ng)
{
    __shared__ double ws[NW];
    group_sums(bsum, nb, gsum, ng, (long long)blockIdx.x * STZ_GRP_PER_BLOCK, ws);
}

// chunk totals of the group sums (one workgroup per chunk of NT groups)
__global__ __launch_bounds__(NT) void k_chunks(const double *__restrict__ gsum, long long ng, double *__restrict__ csum)
{
    __shared__ double sh[NW];
    const double t = chunk_total(gsum, ng, blockIdx.x, sh);
    if (threadIdx.x == NT - 1) csum[blockIdx.x] = t;
}

// One workgroup: R from the chunk totals, the Q<0 count and, when `decide` is set, the
// step kind (STZ/bmg.c:161, STZ/evolution.c:134-136), the selection target and, for an
// athermal step, the list offsets of the avalanche (more than `cap` voxels: overflow).
__global__ __launch_bounds__(NT) void k_decide(const double *__restrict__ csum, long long ng, double *__restrict__ res,
                                               int decide, double dt, const double *dicep, const int *acnt, int *aoff,
                                               long long nblk, long long cap)
{
    __shared__ unsigned long long s_nn;
    __shared__ double s_R;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const double Rt = chunks_total(csum, (ng + NT - 1) / NT);
        unsigned long long *cnt = reinterpret_cast<unsigned long long *>(res + RES_NNEG);
        s_nn = *cnt;
        s_R = Rt;
        *cnt = 0ull;
        res[8] = Rt;
        res[9] = (double)s_nn;
    }
    __syncthreads();
    const double R = s_R;
    if (!decide) return;
    const unsigned long long nn = s_nn;
    int kind;
    long long n = 0;
    if (nn > 0) {
        n = block_offsets(acnt, nblk, aoff);
        kind = n > cap ? KIND_ATH_OVERFLOW : KIND_ATHERMAL;
    } else {
        kind = (1 / R < dt) ? KIND_THERMAL : KIND_ELASTIC;
    }
    if (tid == 0) {
        const double dice = *dicep;
        if (kind == KIND_THERMAL && !(dice > 0.0)) kind = KIND_FAIL;
        reinterpret_cast<int *>(res + RES_KIND)[0] = kind;
        res[RES_TAU] = dice * R;
        reinterpret_cast<int *>(res + 15)[0] = 0;     // NaN flag of the coming solve
        res[13] = (double)n;
        res[14] = 0.0;                                 // athermal sum |DWork|
    }
}

template <typename R, int DIM, bool XTAL>
__global__ __launch_bounds__(NT) void k_select(RateArgs<R> a, const double *__restrict__ gsum,
                                               const double *__restrict__ csum, long long ng, double tau,
                                               double *__restrict__ res, int gated, ApplyArgs ap,
                                               const double *__restrict__ dnew)
{
    if (gated) {
        if (reinterpret_cast<const int *>(res + RES_KIND)[0] != KIND_THERMAL) return;
        tau = res[RES_TAU];
    }
    const bool ok = select_block<R, DIM, XTAL>(a, gsum, csum, ng, tau, res, ap, dnew, 1.0);
    if (!ok && threadIdx.x == 0) reinterpret_cast<int *>(res + RES_KIND)[0] = KIND_FAIL;
}

// athermal step of the fused sequence: list offsets from k_decide, host-drawn redraws
// for the first `cap` voxels, |DWork| partials per workgroup
template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_athermal(AthArgs<R> a, const double *res)
{
    const int kind = reinterpret_cast<const int *>(res + RES_KIND)[0];
    if (kind != KIND_ATHERMAL && kind != KIND_ATH_OVERFLOW) return;
    athermal_block<R, DIM>(a, blockIdx.x);
}

// fixed-order sum of the |DWork| partials (STZ/evolution.c:66-68)
__global__ __launch_bounds__(STZ_NT_MEAN) void k_socsum(const double *__restrict__ socp, long long nblk, double *res)
{
    const int kind = reinterpret_cast<const int *>(res + RES_KIND)[0];
    if (kind != KIND_ATHERMAL && kind != KIND_ATH_OVERFLOW) return;
    const double s = ordered_sum<STZ_NT_MEAN>(socp, nblk);
    if (threadIdx.x == 0) res[14] = s;
}

template <typename R, int DIM>
__global__ void k_apply_one(ApplyArgs a, long long v, int mm, const double *__restrict__ dnew, const double *res)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (res) {      // fused step: the event comes from k_select, only on a thermal step
        if (reinterpret_cast<const int *>(res + RES_KIND)[0] != KIND_THERMAL) return;
        v = (long long)res[10];
        mm = (int)res[11];
    }
    transform_voxel<R, DIM>(a, v, mm, dnew, 1.0);
}

// athermal: exclusive offsets of the per-block counts written by k_rates
// (three phases: per-thread runs, one block scan, per-thread writes)
constexpr int NT_OFF = 1024;
__global__ __launch_bounds__(NT_OFF) void k_aoffsets(const int *__restrict__ cnt, long long nblk, int *__restrict__ off,
                                                     double *__restrict__ res)
{
    __shared__ long long sh[NT_OFF];
    const int tid = threadIdx.x;
    const long long per = (nblk + NT_OFF - 1) / NT_OFF;
    const long long b0 = (long long)tid * per;
    const long long b1 = b0 + per < nblk ? b0 + per : nblk;
    long long run = 0;
    for (long long b = b0; b < b1; ++b) run += cnt[b];
    sh[tid] = run;
    __syncthreads();
    for (int d = 1; d < NT_OFF; d <<= 1) {
        const long long y = tid >= d ? sh[tid - d] : 0;
        __syncthreads();
        sh[tid] += y;
        __syncthreads();
    }
    long long base = sh[tid] - run;
    for (long long b = b0; b < b1; ++b) {
        off[b] = (int)base;
        base += cnt[b];
    }
    if (tid == NT_OFF - 1) res[13] = (double)sh[NT_OFF - 1];
}

template <typename R>
__global__ __launch_bounds__(NT) void k_acompact(const R *__restrict__ minq, long long N, double thr,
                                                 const int *__restrict__ off, int *__restrict__ list)
{
    __shared__ int ws[NW];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long v = (long long)blockIdx.x * NT + tid;
    int f = 0;
    if (v < N) {
        const R q = minq[v];
        f = (q < (R)0 || fabs((double)q) < thr);
    }
    const unsigned long long bal = __ballot(f);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(bal);
    __syncthreads();
    int base = off[blockIdx.x];
    for (int i = 0; i < w; ++i) base += ws[i];
    if (f) list[base + before] = (int)v;
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_aapply(ApplyArgs a, const int *__restrict__ list, int n, const uint8_t *__restrict__ argm,
                                               const R *__restrict__ sigma, double pend, int pcomp, const double *soff,
                                               double voxel, const double *__restrict__ dnew, double *__restrict__ socp)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    __shared__ double red[NT];
    const int i = blockIdx.x * NT + threadIdx.x;
    double dw = 0.0;
    if (i < n) {
        const long long v = list[i];
        const int mm = argm[v];
        // DWork of the chosen mode with the current stress (STZ/evolution.c:66-67)
        R sg[C], st[5], d[C];
        for (int c = 0; c < C; ++c) {
            R x = sigma[(long long)c * a.N + v];
            if constexpr (sizeof(R) == 8) { if (c == pcomp) x = x + (R)pend; }
            else x = (R)((double)x + soff[c]);
            sg[c] = x;
        }
        const R *deps = (const R *)a.deps;
        for (int k = 0; k < S; ++k) st[k] = deps[((long long)mm * S + k) * a.N + v];
        expand<R, DIM>(st, d);
        dw = fabs((double)(contract<R, DIM>(sg, d) * (R)voxel * (R)1e-18 * (R)STZ_J_IN_EV));
        transform_voxel<R, DIM>(a, v, mm, dnew + (long long)i * a.M * S, 1.0);
    }
    red[threadIdx.x] = dw;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) socp[blockIdx.x] = red[0];
}

// elastic energy partials (ElasticEnergy, STZ/elastic.c:533-547)
#pragma clang fp contract(off)
template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_energy(const R *__restrict__ sigma, long long N, double lam, double mu, int pcomp,
                                               double pend, const double *soff, double *__restrict__ part)
{
    constexpr int C = Comp<DIM>::C;
    __shared__ double red[NT];
    const long long v = (long long)blockIdx.x * NT + threadIdx.x;
    double e = 0.0;
    if (v < N) {
        double s[C], el[C];
        for (int c = 0; c < C; ++c) {
            double x;
            if constexpr (sizeof(R) == 8) { x = sigma[(long long)c * N + v]; if (c == pcomp) x = x + pend; }
            else x = (double)sigma[(long long)c * N + v] + soff[c];
            s[c] = x;
        }
        if constexpr (DIM == 2) {
            double tra = -lam / 2 / mu / (2 * lam + 2 * mu) * (s[0] + s[1]);
            el[0] = tra + s[0] / 2 / mu;
            el[1] = tra + s[1] / 2 / mu;
            el[2] = s[2] / 2 / mu;
        } else {
            double tra = -lam / 2 / mu / (3 * lam + 2 * mu) * ((s[0] + s[1]) + s[2]);
            for (int c = 0; c < 3; ++c) el[c] = tra + s[c] / 2 / mu;
            for (int c = 3; c < 6; ++c) el[c] = s[c] / 2 / mu;
        }
        e = 0.5 * contract<double, DIM>(s, el);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
#pragma clang fp contract(on)

template <typename R>
RateArgs<R> rate_args(stz_ctx *ctx)
{
    RateArgs<R> a;
    a.sigma = (R *)ctx->sigma;
    a.q0 = (const R *)ctx->q0;
    a.deps = (const R *)ctx->deps;
    a.deps_last = (const R *)ctx->deps_last;
    a.q0eff = (R *)ctx->q0eff;
    a.scache = (R *)ctx->scache;
    a.perm = ctx->perm;
    a.temp = ctx->temp;
    a.tlast = ctx->tlast;
    a.minq = (R *)ctx->minq;
    a.mro = ctx->mro;
    a.argm = ctx->argm;
    a.bsum = ctx->bsum;
    a.acnt = ctx->acnt;
    a.nneg = reinterpret_cast<unsigned long long *>(ctx->res + RES_NNEG);
    a.thr = ctx->k.thr;
    a.N = ctx->N;
    a.nblk = ctx->nblk;
    a.M = ctx->M;
    a.softening = ctx->p.softening == 1;
    a.pend_comp = ctx->pend_comp;
    a.pend_delta = ctx->pend_delta;
    for (int c = 0; c < 6; ++c) a.soff[c] = ctx->soff[c];
    a.totime = ctx->totime;
    a.tTemp = ctx->k.tTemp;
    a.kT = ctx->k.kT;
    a.ratepre = ctx->k.ratepre;
    a.voxel = ctx->k.voxel;
    a.lnpre = std::log(ctx->k.ratepre);
    return a;
}

ApplyArgs apply_args(stz_ctx *ctx)
{
    ApplyArgs a;
    a.eps0 = ctx->eps0;
    a.deps = ctx->deps;
    a.deps_last = ctx->deps_last;
    a.scache = ctx->scache;
    a.perm = ctx->perm;
    a.temp = ctx->temp;
    a.tlast = ctx->tlast;
    a.tloc = ctx->tloc;
    a.sigma = ctx->sigma;
    for (int c = 0; c < 6; ++c) a.soff[c] = ctx->soff[c];
    // FP64: a uniform increment not yet consumed by a rate pass is not in the field
    if (ctx->prec == STZ_FP64 && ctx->pend_comp >= 0) a.soff[ctx->pend_comp] += ctx->pend_delta;
    a.q0 = ctx->q0;
    a.mro = ctx->mro;
    a.mro_q0 = ctx->p.mro_q0;
    a.rratio = ctx->p.random_ratio;
    a.unew = ctx->dspec_dev + 1 + (size_t)ctx->M * ctx->S;   // pinned: the thermal event's uniforms
    a.N = ctx->N;
    a.M = ctx->M;
    a.permsoft = ctx->k.permsoft;
    a.tempsoft = ctx->k.tempsoft;
    a.permcap = ctx->k.permcap;
    a.totime = ctx->totime;
    return a;
}

template <typename R, int DIM>
int rates_impl(stz_ctx *ctx, int decide)
{
    RateArgs<R> a = rate_args<R>(ctx);
    tmark(ctx, 8);
    if (ctx->mro)
        hipLaunchKernelGGL((k_rates<R, DIM, true>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_rates<R, DIM, false>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 9);
    const long long nb = (long long)ctx->M * ctx->nblk;
    const unsigned gb = (unsigned)((ctx->ngrp + STZ_GRP_PER_BLOCK - 1) / STZ_GRP_PER_BLOCK);
    hipLaunchKernelGGL(k_group, dim3(gb), dim3(NT), 0, ctx->stream, ctx->bsum, nb, ctx->gsum, ctx->ngrp);
    hipLaunchKernelGGL(k_chunks, dim3((unsigned)((ctx->ngrp + NT - 1) / NT)), dim3(NT), 0, ctx->stream, ctx->gsum,
                       ctx->ngrp, ctx->csum);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(NT), 0, ctx->stream, ctx->csum, ctx->ngrp, ctx->res, decide, ctx->k.dt,
                       ctx->dspec_dev, ctx->acnt, ctx->aoff, ctx->nblk, ctx->acap);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 10);
    return STZ_OK;
}

template <typename R, int DIM>
int select_impl(stz_ctx *ctx, double tau, int gated)
{
    RateArgs<R> a = rate_args<R>(ctx);
    a.pend_comp = -1;
    tmark(ctx, 11);
    const double *dnew = gated ? ctx->dspec_dev + 1 : nullptr;
    if (ctx->mro)
        hipLaunchKernelGGL((k_select<R, DIM, true>), dim3(1), dim3(NT), 0, ctx->stream, a, ctx->gsum, ctx->csum, ctx->ngrp, tau,
                           ctx->res, gated, apply_args(ctx), dnew);
    else
        hipLaunchKernelGGL((k_select<R, DIM, false>), dim3(1), dim3(NT), 0, ctx->stream, a, ctx->gsum, ctx->csum, ctx->ngrp, tau,
                           ctx->res, gated, apply_args(ctx), dnew);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 12);
    return STZ_OK;
}

}  // namespace

int stz_rates_launch(stz_ctx *ctx, int decide)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? rates_impl<double, 2>(ctx, decide) : rates_impl<double, 3>(ctx, decide);
    return ctx->dim == 2 ? rates_impl<float, 2>(ctx, decide) : rates_impl<float, 3>(ctx, decide);
}

int stz_select_launch(stz_ctx *ctx, double target, int gated)
{
    if (ctx->prec == STZ_FP64)
        return ctx->dim == 2 ? select_impl<double, 2>(ctx, target, gated) : select_impl<double, 3>(ctx, target, gated);
    return ctx->dim == 2 ? select_impl<float, 2>(ctx, target, gated) : select_impl<float, 3>(ctx, target, gated);
}

int stz_apply_launch(stz_ctx *ctx, int site, int mode, int gated)
{
    ApplyArgs a = apply_args(ctx);
    const double *res = gated ? ctx->res : nullptr;
    const double *dnew = gated ? ctx->dspec_dev + 1 : ctx->dnew;
    const long long v = site;
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<double, 2>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
        else hipLaunchKernelGGL((k_apply_one<double, 3>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
    } else {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<float, 2>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
        else hipLaunchKernelGGL((k_apply_one<float, 3>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_athermal_launch(stz_ctx *ctx)
{
    const double thr = ctx->k.thr;
    const dim3 g((unsigned)ctx->nblk);
    tmark(ctx, 15);
    hipLaunchKernelGGL(k_aoffsets, dim3(1), dim3(NT_OFF), 0, ctx->stream, ctx->acnt, ctx->nblk, ctx->aoff, ctx->res);
    STZ_CHECK_LAUNCH();
    if (ctx->prec == STZ_FP64)
        hipLaunchKernelGGL((k_acompact<double>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    else
        hipLaunchKernelGGL((k_acompact<float>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 16);
    return STZ_OK;
}

template <typename R, int DIM>
int athermal_fused_impl(stz_ctx *ctx, int gated)
{
    AthArgs<R> a;
    a.ap = apply_args(ctx);
    a.minq = (const R *)ctx->minq;
    a.argm = ctx->argm;
    a.sigma = (const R *)ctx->sigma;
    for (int c = 0; c < 6; ++c) a.soff[c] = ctx->soff[c];
    a.voxel = ctx->k.voxel;
    a.thr = ctx->k.thr;
    a.list = ctx->alist;
    a.off = ctx->aoff;
    a.dsrc = ctx->dath_dev;
    a.dscale = 1.0;
    a.cap = ctx->acap;
    a.N = ctx->N;
    a.socp = ctx->socp;
    (void)gated;
    tmark(ctx, 15);
    hipLaunchKernelGGL((k_athermal<R, DIM>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a, ctx->res);
    hipLaunchKernelGGL(k_socsum, dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->socp, ctx->nblk, ctx->res);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 16);
    return STZ_OK;
}

int stz_athermal_fused_launch(stz_ctx *ctx, int gated)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? athermal_fused_impl<double, 2>(ctx, gated) : athermal_fused_impl<double, 3>(ctx, gated);
    return ctx->dim == 2 ? athermal_fused_impl<float, 2>(ctx, gated) : athermal_fused_impl<float, 3>(ctx, gated);
}

int stz_athermal_apply_launch(stz_ctx *ctx, int n) { return stz_athermal_apply_range(ctx, 0, n); }

int stz_athermal_apply_range(stz_ctx *ctx, long long first, long long count)
{
    const int n = (int)count;
    if (n <= 0) return STZ_OK;
    ApplyArgs a = apply_args(ctx);
    const int *alist = ctx->alist + first;
    const dim3 g((unsigned)((n + NT - 1) / NT));
    double *soff = ctx->res + 16;   // device copy of the mixed-precision offsets
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<double, 2>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<double, 3>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<float, 2>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const float *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<float, 3>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const float *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_energy_launch(stz_ctx *ctx)
{
    const dim3 g((unsigned)ctx->nblk);
    double *soff = ctx->res + 16;
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_energy<double, 2>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
        else
            hipLaunchKernelGGL((k_energy<double, 3>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_energy<float, 2>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
        else
            hipLaunchKernelGGL((k_energy<float, 3>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}
