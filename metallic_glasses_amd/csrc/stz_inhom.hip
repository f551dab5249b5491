// stz_inhom.hip -- stress solve of an elastically inhomogeneous glass.
//
// The reference compiles InhomogeneousStressSolver (STZ/elastic.c:977-1066) with
// per-voxel moduli Lambda(x), Mu(x) (STZ/init.c:440-465; its input sets them uniform).
// With moduli set by stz_set_moduli, a stress solve iterates the reference's fixed point
// of the reference-medium eigenstrain eps*:
//
//   Tau0 = C(x):eps0,  eps* = C0^-1 (Tau0 + (C0 - C(x)):(eps0 + epsavg))     (:993-1009)
//   repeat:  epsolast = eps*;  homogeneous solve of epsolast -> total strain eps, sigma
//            eps* = 0.5 C0^-1 (Tau0 + (C0 - C(x)):eps) + 0.5 epsolast        (:1017-1046)
//   until sqrt(sum_c sum_v (eps* - epsolast)^2 / N) < 1e-3 |epsavg| or |epsavg| < 1e-4,
//   or after 101 homogeneous solves                                           (:1048-1062)
//
// Each iteration is one homogeneous solve whose C2R x pass also takes the update: the
// epilogue has the total strain, the eigenstrain it came from, Tau0 and the voxel's
// moduli in registers, so eps* is written and its squared change summed per block in the
// same pass (no extra N-passes over the fields).  The host reads one double per
// iteration to decide whether the loop ends.  |epsavg| and the change run over the C
// stored components (the reference's three in 2-D).
#include "stz_solve_dev.h"

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr int NT_INIT = 256;

#pragma clang fp contract(off)
// Tau0 and the first eps* (IsotropicMul_RF, IsotropicMul_RF_D, IsotropicInv on
// eps0 and eps0 + epsavg, STZ/elastic.c:993-1009)
template <int C>
__global__ __launch_bounds__(NT_INIT) void k_inhom_init(const double *__restrict__ eps0,
                                                        const double *__restrict__ lamx,
                                                        const double *__restrict__ mux, double *__restrict__ tau0,
                                                        double *__restrict__ eout, SolveConsts k)
{
    const long long v = (long long)blockIdx.x * NT_INIT + threadIdx.x;
    if (v >= k.N) return;
    double e[C], z[C], t0[C], t1[C], tau[C], r[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        e[c] = eps0[(long long)c * k.N + v];
        z[c] = e[c] + k.epsavg[c];
    }
    const double lx = lamx[v], mx = mux[v];
    iso_mul<C>(lx, mx, e, t0);
    iso_mul<C>(k.lambda - lx, k.mu - mx, z, t1);
#pragma unroll
    for (int c = 0; c < C; ++c) tau[c] = t0[c] + t1[c];
    iso_inv<C>(k.lambda, k.mu, tau, r);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        tau0[(long long)c * k.N + v] = t0[c];
        eout[(long long)c * k.N + v] = r[c];
    }
}
#pragma clang fp contract(on)

// fixed-order sums of per-block partials: C stress components (scaled by invN) or the
// squared eigenstrain change (C = 1, invN = 1)
template <int C>
__global__ __launch_bounds__(STZ_NT_MEAN) void k_partials(const double *__restrict__ part, int nblocks, double invN,
                                                          double *__restrict__ out)
{
    mean_partials<C>(part, nblocks, invN, out);
}

// One x pass of the iteration for axis length L0 with the uniform solve's default row
// count: the largest power of two up to rows_for(L0) that divides the rows (so the
// stress partial sums, and <sigma>, group as the uniform solve's do).
template <typename R, int C, int L0>
int x_pass(stz_ctx *ctx, const SolveConsts &k, bool inverse, const double *src, int *nblocks)
{
    constexpr int NRD = rows_for<R, C>(L0);
    auto go = [&](auto NRc) -> int {
        constexpr int NRV = decltype(NRc)::value;
        if constexpr (x_variant<R, C>(L0, NRV)) {
            const dim3 grid((unsigned)(k.rows / NRV));
            if (nblocks) *nblocks = k.rows / NRV;
            if (!inverse)
                hipLaunchKernelGGL((k_fwd_x<R, L0, NRV, C, 1>), grid, dim3(NT_SOLVE), 0, ctx->stream, src,
                                   (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[0], k, (const SolveRep *)nullptr);
            else
                hipLaunchKernelGGL((k_inv_x<R, L0, NRV, C, 1>), grid, dim3(NT_SOLVE), 0, ctx->stream,
                                   (cpx<R> *)ctx->spec, src, (R *)ctx->sigma, ctx->spart,
                                   reinterpret_cast<int *>(ctx->res + 15), (const cpx<R> *)ctx->tw.tw[0], k,
                                   (const SolveRep *)nullptr);
            STZ_CHECK_LAUNCH();
            return STZ_OK;
        } else {
            ctx->err = "no compiled x-pass variant for this grid";
            return STZ_E_UNSUPPORTED;
        }
    };
    int NR = NRD;
    while (NR > 2 && k.rows % NR != 0) NR /= 2;
    if (k.rows % NR != 0) {
        ctx->err = "the row count of this grid is odd";
        return STZ_E_ARG;
    }
    switch (NR) {
    case 32: if constexpr (NRD >= 32) return go(std::integral_constant<int, 32>()); break;
    case 16: if constexpr (NRD >= 16) return go(std::integral_constant<int, 16>()); break;
    case 8: if constexpr (NRD >= 8) return go(std::integral_constant<int, 8>()); break;
    case 4: if constexpr (NRD >= 4) return go(std::integral_constant<int, 4>()); break;
    case 2: return go(std::integral_constant<int, 2>());
    default: break;
    }
    ctx->err = "no compiled x-pass variant for this grid";
    return STZ_E_UNSUPPORTED;
}

template <typename R>
int x_dispatch(stz_ctx *ctx, const SolveConsts &k, bool inverse, const double *src, int *nblocks)
{
    int rc = STZ_OK;
    if (ctx->C == 3) {
        STZ_LEN_DISPATCH(ctx->n[0], rc = (x_pass<R, 3, LL>(ctx, k, inverse, src, nblocks)));
    } else {
        STZ_LEN_DISPATCH(ctx->n[0], rc = (x_pass<R, 6, LL>(ctx, k, inverse, src, nblocks)));
    }
    return rc;
}

// one homogeneous solve of estar[cur] with the update into estar[cur ^ 1]
template <typename R>
int iterate(stz_ctx *ctx, int cur, int *nblocks)
{
    SolveConsts k = solve_consts(ctx);
    k.tau0 = ctx->tau0;
    k.lamx = ctx->lamx;
    k.mux = ctx->mux;
    k.eout = ctx->estar[cur ^ 1];
    k.cpart = ctx->cpart;
    int rc = x_dispatch<R>(ctx, k, false, ctx->estar[cur], nullptr);
    if (rc) return rc;
    rc = stz_solve_spectral(ctx);
    if (rc) return rc;
    return x_dispatch<R>(ctx, k, true, ctx->estar[cur], nblocks);
}

void free_all(stz_ctx *ctx)
{
    double **ptrs[] = {&ctx->lamx, &ctx->mux, &ctx->tau0, &ctx->estar[0], &ctx->estar[1], &ctx->cpart};
    for (double **p : ptrs) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

}  // namespace

int stz_inhom_fwd_x(stz_ctx *ctx, const double *src)
{
    const SolveConsts k = solve_consts(ctx);
    return ctx->prec == STZ_FP64 ? x_dispatch<double>(ctx, k, false, src, nullptr)
                                 : x_dispatch<float>(ctx, k, false, src, nullptr);
}

int stz_inhom_run(stz_ctx *ctx)
{
    const long long N = ctx->N;
    const int C = ctx->C;
    const SolveConsts k = solve_consts(ctx);
    const dim3 g((unsigned)((N + NT_INIT - 1) / NT_INIT));
    tmark(ctx, 0);
    if (C == 3)
        hipLaunchKernelGGL((k_inhom_init<3>), g, dim3(NT_INIT), 0, ctx->stream, ctx->eps0, ctx->lamx, ctx->mux,
                           ctx->tau0, ctx->estar[0], k);
    else
        hipLaunchKernelGGL((k_inhom_init<6>), g, dim3(NT_INIT), 0, ctx->stream, ctx->eps0, ctx->lamx, ctx->mux,
                           ctx->tau0, ctx->estar[0], k);
    STZ_CHECK_LAUNCH();
    double ne = 0.0;
    for (int c = 0; c < C; ++c) ne += ctx->epsavg[c] * ctx->epsavg[c];
    const double normepsavg = std::sqrt(ne);          // STZ/elastic.c:1012
    int cur = 0, step = 0, nblocks = 0;
    for (;;) {
        ++step;
        int rc = ctx->prec == STZ_FP64 ? iterate<double>(ctx, cur, &nblocks) : iterate<float>(ctx, cur, &nblocks);
        if (rc) return rc;
        // |epsavg| < 1e-4 ends the loop whatever the change (:1049), and so does step > 100 (:1057)
        if (normepsavg < 1E-4 || step > 100) break;
        hipLaunchKernelGGL((k_partials<1>), dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->cpart, nblocks, 1.0,
                           ctx->res + RES_CONV);
        STZ_CHECK_LAUNCH();
        STZ_HIP(hipMemcpyAsync(ctx->hres + RES_CONV, ctx->res + RES_CONV, sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        STZ_HIP(hipStreamSynchronize(ctx->stream));
        const double convg = std::sqrt(ctx->hres[RES_CONV] / (double)N);   // AvgChange, :957-975
        if (convg < (1E-3) * normepsavg) break;
        cur ^= 1;
    }
    ctx->inhom_cur = cur;                             // epsolast of the last iteration
    ctx->inhom_iters = step;
    if (C == 3)
        hipLaunchKernelGGL((k_partials<3>), dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->spart, nblocks, k.invN,
                           ctx->res);
    else
        hipLaunchKernelGGL((k_partials<6>), dim3(1), dim3(STZ_NT_MEAN), 0, ctx->stream, ctx->spart, nblocks, k.invN,
                           ctx->res);
    STZ_CHECK_LAUNCH();
    for (int i = 1; i <= 6; ++i) tmark(ctx, i);      // the whole loop counts as one solve
    return STZ_OK;
}

int stz_inhom_free(stz_ctx *ctx)
{
    free_all(ctx);
    ctx->inhom_cur = 0;
    ctx->inhom_iters = 1;
    return STZ_OK;
}

extern "C" {

int stz_set_moduli(stz_ctx *ctx, const double *lambda, const double *mu)
{
    if (!ctx || (lambda == nullptr) != (mu == nullptr)) return STZ_E_ARG;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return STZ_E_HIP;
    if (!lambda) {
        stz_inhom_free(ctx);
        ctx->tune_epoch++;
        return STZ_OK;
    }
    if (ctx->solver_mode == STZ_SOLVER_LOCAL) {
        ctx->err = "the event-local stress update assumes uniform moduli";
        return STZ_E_UNSUPPORTED;
    }
    const size_t N = (size_t)ctx->N;
    for (size_t v = 0; v < N; ++v)
        if (!std::isfinite(lambda[v]) || !std::isfinite(mu[v])) {
            ctx->err = "non-finite modulus";
            return STZ_E_ARG;
        }
    if (!ctx->lamx) {
        const size_t rows = (size_t)ctx->n[1] * ctx->n[2];
        const size_t sizes[6] = {N, N, (size_t)ctx->C * N, (size_t)ctx->C * N, (size_t)ctx->C * N, rows};
        double **ptrs[6] = {&ctx->lamx, &ctx->mux, &ctx->tau0, &ctx->estar[0], &ctx->estar[1], &ctx->cpart};
        for (int i = 0; i < 6; ++i)
            if (hipMalloc((void **)ptrs[i], sizes[i] * sizeof(double)) != hipSuccess) {
                free_all(ctx);
                ctx->err = "out of device memory for the inhomogeneous solve";
                return STZ_E_HIP;
            }
    }
    STZ_HIP(hipMemcpyAsync(ctx->lamx, lambda, N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    STZ_HIP(hipMemcpyAsync(ctx->mux, mu, N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx->inhom_cur = 0;
    ctx->tune_epoch++;
    return STZ_OK;
}

int stz_inhom_iterations(stz_ctx *ctx) { return ctx ? ctx->inhom_iters : STZ_E_ARG; }

}  // extern "C"
