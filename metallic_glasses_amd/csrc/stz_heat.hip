// stz_heat.hip -- HeatConducting (STZ/elastic.c:549-582): spectral diffusion of the local
// temperature field.
//
// The reference (flag HEAT_CONDUCT) transforms Tlocal, multiplies its spectrum by
// exp(-g2*dt) -- g2 = |k| of its k grid (STZ/init.c:116-137), 1 at k = 0 -- and
// transforms back, after every thermal step (dt = thermal_diffusivity * t_kmc,
// STZ/evolution.c:247-249) and every elastic step (dt = thermal_diffusivity * dt,
// :325-327).  Tlocal itself is heated by every transform, |sigma:deps| * 1e9 / (Cp rho)
// (STZ/elastic.c:381-383), with or without the flag; that part lives in the transform.
//
// One real component through the stress solve's passes: the R2C x pass (raw input, no
// round trip), the y passes (3-D), the fused last-axis pass with the multiplier in place
// of Gamma-hat (OP 2), and a C2R x pass that writes Tlocal = Re/N in FP64.  The
// multiplier is even in k, so the half spectrum reproduces Re() of the full complex
// transform without the Nyquist averaging the stress operator needs.
#include "stz_solve_dev.h"

#include <cmath>
#include <type_traits>

namespace {

// C2R along x of the single component, Tlocal = Re(IFFT)/N (STZ/elastic.c:576-579)
template <typename R, int L0, int NR>
__global__ __launch_bounds__(NT_SOLVE) void k_inv_x_heat(const cpx<R> *__restrict__ spec, double *__restrict__ out,
                                                         int *__restrict__ nanflag, const cpx<R> *__restrict__ twg,
                                                         SolveConsts k)
{
    constexpr int NP = NR / 2;
    constexpr int LS = NP + 1;
    constexpr int H = L0 / 2;
    __shared__ cpx<R> buf[L0 * LS];
    __shared__ cpx<R> tw[L0];
    const int tid = threadIdx.x;
    for (int i = tid; i < L0; i += NT_SOLVE) tw[i] = twg[i];
    const long long row0 = (long long)blockIdx.x * NR;
    for (int e = tid; e < NP * L0; e += NT_SOLVE) {
        const int kk = e % L0;
        const int p = e / L0;
        const cpx<R> *ra = spec + (row0 + 2 * p) * k.hp;
        const cpx<R> *rb = ra + k.hp;
        cpx<R> Z;
        if (kk == 0 || kk == H) {
            Z = {ra[kk].x, rb[kk].x};
        } else if (kk < H) {
            const cpx<R> a = ra[kk], b = rb[kk];
            Z = {a.x - b.y, a.y + b.x};
        } else {
            const cpx<R> a = ra[L0 - kk], b = rb[L0 - kk];
            Z = {a.x + b.y, -a.y + b.x};
        }
        buf[kk * LS + p] = Z;
    }
    __syncthreads();
    fft_lines<R, L0, NP, LS, NT_SOLVE>(buf, tw, (R)1);
    const R *bufr = reinterpret_cast<const R *>(buf);
    int bad = 0;
    for (int e = tid; e < NR * L0; e += NT_SOLVE) {
        const int rl = e / L0, i0 = e % L0;
        const long long v = (row0 + rl) * L0 + i0;
        const double z = (double)bufr[2 * (i0 * LS + (rl >> 1)) + (rl & 1)];
        bad |= (z != z);
        out[v] = z / (double)k.N;
    }
    if (bad) atomicOr(nanflag, 1);
}

template <typename R, int L0>
int x_heat(stz_ctx *ctx, const SolveConsts &k, bool inverse)
{
    constexpr int NRD = rows_for<R, 1>(L0);
    int NR = NRD;
    while (NR > 2 && k.rows % NR != 0) NR /= 2;
    if (k.rows % NR != 0) {
        ctx->err = "the row count of this grid is odd";
        return STZ_E_ARG;
    }
    auto go = [&](auto NRc) -> int {
        constexpr int NRV = decltype(NRc)::value;
        if constexpr (x_variant<R, 1>(L0, NRV)) {
            const dim3 grid((unsigned)(k.rows / NRV));
            if (!inverse)
                hipLaunchKernelGGL((k_fwd_x<R, L0, NRV, 1, 1>), grid, dim3(NT_SOLVE), 0, ctx->stream, ctx->tloc,
                                   (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[0], k, (const SolveRep *)nullptr);
            else
                hipLaunchKernelGGL((k_inv_x_heat<R, L0, NRV>), grid, dim3(NT_SOLVE), 0, ctx->stream,
                                   (const cpx<R> *)ctx->spec, ctx->tloc, reinterpret_cast<int *>(ctx->res + 15),
                                   (const cpx<R> *)ctx->tw.tw[0], k);
            STZ_CHECK_LAUNCH();
            return STZ_OK;
        } else {
            ctx->err = "no compiled heat x-pass variant for this grid";
            return STZ_E_UNSUPPORTED;
        }
    };
    switch (NR) {
    case 32: if constexpr (NRD >= 32) return go(std::integral_constant<int, 32>()); break;
    case 16: if constexpr (NRD >= 16) return go(std::integral_constant<int, 16>()); break;
    case 8: if constexpr (NRD >= 8) return go(std::integral_constant<int, 8>()); break;
    case 4: if constexpr (NRD >= 4) return go(std::integral_constant<int, 4>()); break;
    case 2: return go(std::integral_constant<int, 2>());
    default: break;
    }
    ctx->err = "no compiled heat x-pass variant for this grid";
    return STZ_E_UNSUPPORTED;
}

template <typename R, int L1>
int y_heat(stz_ctx *ctx, const SolveConsts &k, R sgn)
{
    constexpr int W = cols_for(L1);
    const dim3 grid((unsigned)((k.h0 + W - 1) / W), (unsigned)k.n[2], 1);
    hipLaunchKernelGGL((k_ypass<R, L1, W>), grid, dim3(NT_SOLVE), 0, ctx->stream, (cpx<R> *)ctx->spec,
                       (const cpx<R> *)ctx->tw.tw[1], k, sgn, (const SolveRep *)nullptr);
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

template <typename R, int LZ, int DIM>
int fused_heat(stz_ctx *ctx, const SolveConsts &k)
{
    constexpr int W = fused_cols_for<R, 1>(LZ);
    constexpr int NTF = fused_threads(LZ * W);
    const dim3 grid((unsigned)((k.h0 + W - 1) / W), DIM == 3 ? (unsigned)k.n[1] : 1u, 1);
    hipLaunchKernelGGL((k_fused<R, LZ, W, 1, DIM, NTF, 2>), grid, dim3(NTF), 0, ctx->stream, (cpx<R> *)ctx->spec,
                       (const cpx<R> *)ctx->tw.tw[DIM - 1], k, (const SolveRep *)nullptr);
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

template <typename R>
int heat_impl(stz_ctx *ctx, double hdt)
{
    SolveConsts k = solve_consts(ctx);
    k.C = 1;                                  // one component: y-pass grid and strides
    k.hdt = hdt;
    int rc = STZ_OK;
    STZ_LEN_DISPATCH(ctx->n[0], rc = (x_heat<R, LL>(ctx, k, false)));
    if (rc) return rc;
    if (ctx->dim == 3) {
        STZ_LEN_DISPATCH(ctx->n[1], rc = (y_heat<R, LL>(ctx, k, (R)-1)));
        if (rc) return rc;
        STZ_LEN_DISPATCH(ctx->n[2], rc = (fused_heat<R, LL, 3>(ctx, k)));
        if (rc) return rc;
        STZ_LEN_DISPATCH(ctx->n[1], rc = (y_heat<R, LL>(ctx, k, (R)1)));
        if (rc) return rc;
    } else {
        STZ_LEN_DISPATCH(ctx->n[1], rc = (fused_heat<R, LL, 2>(ctx, k)));
        if (rc) return rc;
    }
    STZ_LEN_DISPATCH(ctx->n[0], rc = (x_heat<R, LL>(ctx, k, true)));
    return rc;
}

}  // namespace

// enqueues one HeatConducting(hdt) of ctx->tloc; the NaN flag is res[15]
int stz_heat_launch(stz_ctx *ctx, double hdt)
{
    return ctx->prec == STZ_FP64 ? heat_impl<double>(ctx, hdt) : heat_impl<float>(ctx, hdt);
}
