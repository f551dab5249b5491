// stz_local.hip -- event-local stress update (see stz_local_dev.h): the impulse response,
// built once per (grid, elastic constants, precision) and shared by contexts, and the
// per-call update kernel behind stz_stress_update_local.
#include "stz_local_dev.h"

#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace {

// K[k][c][i] = (R) sigma_k[c][i], i < H: the half space of offsets is the first H voxels
template <typename R>
__global__ void k_keep(const double *__restrict__ sig, R *__restrict__ K, int C, int N, int H)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)C * H) return;
    const int c = (int)(i / H), j = (int)(i % H);
    K[i] = (R)sig[(long long)c * N + j];
}

// the compact table (stz_local_dev.h, Resp) from the S full responses F[k][c][i] (FP64):
// A_ik = t_i : sigma_k with the symmetric pair averaged, and the responses of the
// reference normal component
template <typename R, int DIM>
__global__ void k_compact(const double *__restrict__ F, R *__restrict__ K, int H)
{
    using Q = Resp<DIM>;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= H) return;
    auto f = [&](int k, int c) { return F[((long long)k * Q::C + c) * H + j]; };
    // row i of A applied to the response of k: normal rows e_i - e_Z, shear rows 2 e_i
    auto a = [&](int i, int k) {
        return i < Q::NN ? f(k, Q::comp(i)) - f(k, Q::Z) : 2.0 * f(k, Q::comp(i));
    };
#pragma unroll
    for (int i = 0; i < Q::S; ++i)
#pragma unroll
        for (int k = i; k < Q::S; ++k) K[(long long)Q::sym(i, k) * H + j] = (R)(0.5 * (a(i, k) + a(k, i)));
#pragma unroll
    for (int k = 0; k < Q::S; ++k) K[(long long)(Q::NSYM + k) * H + j] = (R)f(k, Q::Z);
}

struct Uni {
    double u[6];
};

template <typename R, int DIM>
__global__ __launch_bounds__(NT_LOCAL) void k_local(LocalGeom g, const R *__restrict__ K, const int *__restrict__ sites,
                                                    int ns, const R *__restrict__ deps_last, R *__restrict__ sigma,
                                                    Uni uni, double *__restrict__ spart, int *nanflag)
{
    __shared__ LocalSrc<R> src[STZ_LOCAL_MAXSRC];
    __shared__ double wred[NW_LOCAL * 6];
    for (int s = threadIdx.x; s < ns; s += NT_LOCAL) load_src<R, DIM>(g, deps_last, sites[s], src[s]);
    __syncthreads();
    local_rows<R, DIM>(g, K, src, ns, uni.u, sigma, blockIdx.x, spart, nanflag, wred);
}

template <int C>
__global__ __launch_bounds__(STZ_NT_MEAN) void k_local_mean(const double *__restrict__ spart, int nblocks, double invN,
                                                            double *__restrict__ out)
{
    mean_partials<C>(spart, nblocks, invN, out);
}

// basis eigenstrains of the stored components, expanded to C tensor components
void basis_tensor(int dim, int k, double *t)
{
    for (int c = 0; c < 6; ++c) t[c] = 0.0;
    if (dim == 2) {
        if (k == 0) t[0] = 1.0, t[1] = -1.0;                // yy = -xx
        else t[2] = 1.0;
    } else {
        if (k == 0) t[0] = 1.0, t[2] = -1.0;                // zz = -(xx + yy)
        else if (k == 1) t[1] = 1.0, t[2] = -1.0;
        else t[k + 1] = 1.0;                                // yz, xz, xy
    }
}

struct KEntry {
    void *k = nullptr;
    int refs = 0;
};
using KKey = std::tuple<int, int, int, int, int, double, double>;
std::mutex g_kmu;
std::map<KKey, KEntry> g_kcache;

KKey key_of(const stz_ctx *ctx)
{
    return KKey(ctx->dim, ctx->n[0], ctx->n[1], ctx->n[2], ctx->prec, ctx->k.lambda, ctx->k.mu);
}

LocalGeom geom_of(const stz_ctx *ctx) { return local_geom(ctx->dim, ctx->n, local_row_voxels(ctx->prec == STZ_MIXED)); }

#define LHIP(call)                                                                 \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            rc = STZ_E_HIP;                                                        \
            goto done;                                                             \
        }                                                                          \
    } while (0)

// planes of the compact response table of a context
int resp_planes(const stz_ctx *ctx) { return ctx->dim == 2 ? Resp<2>::P : Resp<3>::P; }

// S FP64 impulse solves on a solver-only shadow context of the same grid and constants
// (the stress-solve kernels read only the arrays allocated here), their half-space
// responses kept in FP64 and compacted into K (resp_planes x H values)
int build_response(stz_ctx *ctx, void *K)
{
    const LocalGeom g = geom_of(ctx);
    const int C = ctx->C;
    const size_t N = (size_t)ctx->N, rows = (size_t)ctx->n[1] * ctx->n[2];
    stz_ctx *sh = new stz_ctx();
    double *full = nullptr;       // [S][C][H] FP64 responses
    int rc = STZ_OK;
    sh->p = ctx->p;
    sh->p.precision = STZ_FP64;
    sh->k = ctx->k;
    sh->C = ctx->C;
    sh->S = ctx->S;
    sh->M = ctx->M;
    sh->dim = ctx->dim;
    for (int a = 0; a < 3; ++a) sh->n[a] = ctx->n[a];
    sh->N = ctx->N;
    sh->lc = ctx->lc;
    sh->prec = STZ_FP64;
    sh->rsz = 8;
    sh->h0 = ctx->h0;
    sh->hp = (ctx->h0 + 7) / 8 * 8;          // 128-byte rows of FP64 complex values
    sh->stream = ctx->stream;
    LHIP(hipMalloc((void **)&sh->eps0, C * N * 8));
    LHIP(hipMalloc(&sh->sigma, C * N * 8));
    LHIP(hipMalloc(&sh->spec, (size_t)C * rows * sh->hp * 16));
    LHIP(hipMalloc((void **)&sh->spart, rows * C * 8));
    LHIP(hipMalloc((void **)&sh->res, 64 * 8));
    LHIP(hipMalloc((void **)&full, (size_t)ctx->S * C * g.H * 8));
    LHIP(hipMemsetAsync(sh->res, 0, 64 * 8, ctx->stream));
    rc = stz_fft_setup(sh);
    if (rc) {
        ctx->err = sh->err;
        goto done;
    }
    for (int k = 0; k < ctx->S; ++k) {
        double t[6];
        basis_tensor(ctx->dim, k, t);
        LHIP(hipMemsetAsync(sh->eps0, 0, C * N * 8, ctx->stream));
        for (int c = 0; c < C; ++c)
            LHIP(hipMemcpyAsync(sh->eps0 + (size_t)c * N, &t[c], 8, hipMemcpyHostToDevice, ctx->stream));
        LHIP(hipStreamSynchronize(ctx->stream));   // t lives on this frame
        rc = stz_solve_launch(sh, 0);
        if (rc) {
            ctx->err = sh->err;
            goto done;
        }
        const long long ch = (long long)C * g.H;
        const unsigned grid = (unsigned)((ch + 255) / 256);
        hipLaunchKernelGGL(k_keep<double>, dim3(grid), dim3(256), 0, ctx->stream, (const double *)sh->sigma,
                           full + (size_t)k * ch, C, g.N, g.H);
        LHIP(hipGetLastError());
    }
    {
        const unsigned grid = (unsigned)((g.H + 255) / 256);
        if (ctx->prec == STZ_FP64 && ctx->dim == 2)
            hipLaunchKernelGGL((k_compact<double, 2>), dim3(grid), dim3(256), 0, ctx->stream, full, (double *)K, g.H);
        else if (ctx->prec == STZ_FP64)
            hipLaunchKernelGGL((k_compact<double, 3>), dim3(grid), dim3(256), 0, ctx->stream, full, (double *)K, g.H);
        else if (ctx->dim == 2)
            hipLaunchKernelGGL((k_compact<float, 2>), dim3(grid), dim3(256), 0, ctx->stream, full, (float *)K, g.H);
        else
            hipLaunchKernelGGL((k_compact<float, 3>), dim3(grid), dim3(256), 0, ctx->stream, full, (float *)K, g.H);
        LHIP(hipGetLastError());
    }
    LHIP(hipStreamSynchronize(ctx->stream));
    {
        int nan = 0;
        LHIP(hipMemcpy(&nan, sh->res + 15, sizeof(int), hipMemcpyDeviceToHost));
        if (nan) {
            ctx->err = "NaN in an impulse-response solve";
            rc = STZ_E_NAN;
        }
    }
done:
    (void)hipStreamSynchronize(ctx->stream);
    void *ptrs[] = {sh->eps0, sh->sigma, sh->spec, sh->spart, sh->res, full};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    stz_fft_free(sh);
    sh->stream = nullptr;
    delete sh;
    return rc;
}

}  // namespace

int stz_local_prepare(stz_ctx *ctx)
{
    if (ctx->kresp) return STZ_OK;
    if (ctx->N >= (1ll << 31)) {
        ctx->err = "the event-local update needs fewer than 2^31 voxels";
        return STZ_E_ARG;
    }
    const LocalGeom g = geom_of(ctx);
    std::lock_guard<std::mutex> lk(g_kmu);
    const KKey key = key_of(ctx);
    auto it = g_kcache.find(key);
    if (it != g_kcache.end()) {
        it->second.refs++;
        ctx->kresp = it->second.k;
    } else {
        void *K = nullptr;
        if (hipMalloc(&K, (size_t)resp_planes(ctx) * g.H * ctx->rsz) != hipSuccess) {
            ctx->err = "hipMalloc of the impulse response failed";
            return STZ_E_HIP;
        }
        const int rc = build_response(ctx, K);
        if (rc) {
            (void)hipFree(K);
            return rc;
        }
        KEntry e;
        e.k = K;
        e.refs = 1;
        g_kcache[key] = e;
        ctx->kresp = K;
    }
    if (!ctx->lsrc && hipMalloc((void **)&ctx->lsrc, STZ_LOCAL_MAXSRC * sizeof(int)) != hipSuccess) {
        ctx->err = "hipMalloc failed";
        return STZ_E_HIP;
    }
    return STZ_OK;
}

void stz_local_release(stz_ctx *ctx)
{
    if (ctx->lsrc) (void)hipFree(ctx->lsrc);
    ctx->lsrc = nullptr;
    if (!ctx->kresp) return;
    std::lock_guard<std::mutex> lk(g_kmu);
    for (auto it = g_kcache.begin(); it != g_kcache.end(); ++it) {
        if (it->second.k == ctx->kresp) {
            if (--it->second.refs == 0) {
                (void)hipFree(it->second.k);
                g_kcache.erase(it);
            }
            break;
        }
    }
    ctx->kresp = nullptr;
}

void stz_local_mark_solved(stz_ctx *ctx)
{
    for (int c = 0; c < 6; ++c) {
        ctx->ebase[c] = ctx->epsavg[c];
        ctx->usum[c] = 0.0;
    }
    ctx->nlocal = 0;
}

// the uniform part of the update: C0:(epsavg - ebase), minus what was already added to
// the field (FP64)
void stz_local_uniform(const stz_ctx *ctx, double *uni)
{
    double de[6];
    for (int c = 0; c < 6; ++c) {
        de[c] = ctx->epsavg[c] - ctx->ebase[c];
        uni[c] = 0.0;
    }
    if (ctx->C == 3) c0_mul<3>(de, ctx->k.lambda, ctx->k.mu, uni);
    else c0_mul<6>(de, ctx->k.lambda, ctx->k.mu, uni);
    if (ctx->prec == STZ_FP64)
        for (int c = 0; c < 6; ++c) uni[c] -= ctx->usum[c];
}

// enqueues the update of the sites already in ctx->lsrc and the mean into ctx->res[0..C)
int stz_local_launch(stz_ctx *ctx, int ns, const double *uni)
{
    LocalGeom g = geom_of(ctx);
    const long long nb = stz_local_blocks(ctx);
    Uni u;
    for (int c = 0; c < 6; ++c) u.u[c] = uni[c];
    int *nanflag = reinterpret_cast<int *>(ctx->res + 15);
    const dim3 grid((unsigned)nb);
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_local<double, 2>), grid, dim3(NT_LOCAL), 0, ctx->stream, g, (const double *)ctx->kresp,
                               ctx->lsrc, ns, (const double *)ctx->deps_last, (double *)ctx->sigma, u, ctx->spart,
                               nanflag);
        else
            hipLaunchKernelGGL((k_local<double, 3>), grid, dim3(NT_LOCAL), 0, ctx->stream, g, (const double *)ctx->kresp,
                               ctx->lsrc, ns, (const double *)ctx->deps_last, (double *)ctx->sigma, u, ctx->spart,
                               nanflag);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_local<float, 2>), grid, dim3(NT_LOCAL), 0, ctx->stream, g, (const float *)ctx->kresp,
                               ctx->lsrc, ns, (const float *)ctx->deps_last, (float *)ctx->sigma, u, ctx->spart,
                               nanflag);
        else
            hipLaunchKernelGGL((k_local<float, 3>), grid, dim3(NT_LOCAL), 0, ctx->stream, g, (const float *)ctx->kresp,
                               ctx->lsrc, ns, (const float *)ctx->deps_last, (float *)ctx->sigma, u, ctx->spart,
                               nanflag);
    }
    STZ_CHECK_LAUNCH();
    if (ctx->C == 3)
        hipLaunchKernelGGL((k_local_mean<3>), dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->spart, (int)nb,
                           1.0 / (double)ctx->N, ctx->res);
    else
        hipLaunchKernelGGL((k_local_mean<6>), dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->spart, (int)nb,
                           1.0 / (double)ctx->N, ctx->res);
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

// workgroups of the update over one grid (partials the mean reduces): one per ipw offset
// rows of the half space, at most n1*(n2/2 + 1) <= n1*n2, the rows of the partial array
long long stz_local_blocks(const stz_ctx *ctx)
{
    const LocalGeom g = geom_of(ctx);
    return (g.items + g.ipw - 1) / g.ipw;
}
