// stz_solve_launch.h -- the host-side launch sequence of the stress solve (uniform moduli)
// and of the print-step displacement, for one storage precision: included by one
// translation unit per precision with STZ_SOLVE_REAL defined (double or float), so that
// the two sets of kernel variants compile in parallel.  The kernels and the data flow are
// described in stz_solve.hip.
#pragma once
#include "stz_solve_dev.h"

#include <cmath>
#include <type_traits>

#ifndef STZ_SOLVE_REAL
#error "define STZ_SOLVE_REAL (double or float) before including stz_solve_launch.h"
#endif

namespace {

// ---------------------------------------------------------------------------------
// k_inv_x_disp: C2R along x of the DIM displacement components left in the component
// slots 0..DIM-1 by the OP 1 fused pass, u_a = Re/N + supercell position
// (STZ/elastic.c:884-899), into u[a][v] (FP64)
// ---------------------------------------------------------------------------------
template <typename R, int L0, int NR, int DIM>
__global__ __launch_bounds__(NT_SOLVE) void k_inv_x_disp(const cpx<R> *__restrict__ spec,
                                                         double *__restrict__ u, const cpx<R> *__restrict__ twg,
                                                         SolveConsts k)
{
    constexpr int NP = NR / 2;
    constexpr int NL = DIM * NP;
    constexpr int LS = NL + 1;
    constexpr int H = L0 / 2;
    __shared__ cpx<R> buf[L0 * LS];
    __shared__ cpx<R> tw[L0];
    const int tid = threadIdx.x;
    for (int i = tid; i < L0; i += NT_SOLVE) tw[i] = twg[i];
    const long long row0 = (long long)blockIdx.x * NR;
    for (int e = tid; e < DIM * NP * L0; e += NT_SOLVE) {
        const int kk = e % L0;
        const int p = (e / L0) % NP;
        const int c = e / (L0 * NP);
        const cpx<R> *ra = spec + ((long long)c * k.rows + row0 + 2 * p) * k.hp;
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
        buf[kk * LS + c * NP + p] = Z;
    }
    __syncthreads();
    fft_lines<R, L0, NL, LS, NT_SOLVE>(buf, tw, (R)1);
    const R *bufr = reinterpret_cast<const R *>(buf);
    for (int e = tid; e < NR * L0; e += NT_SOLVE) {
        const int rl = e / L0, i0 = e % L0;
        const long long r = row0 + rl;
        const long long v = r * L0 + i0;
        const int i1 = (int)(r % k.n[1]), i2 = (int)(r / k.n[1]);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            const double z = (double)bufr[2 * (i0 * LS + a * NP + (rl >> 1)) + (rl & 1)];
            u[(long long)a * k.N + v] = z / (double)k.N + position<DIM>(a, i0, i1, i2, k.epsavg, k.L);
        }
    }
}

constexpr int NT_MEAN = STZ_NT_MEAN;
template <int C>
__global__ __launch_bounds__(NT_MEAN) void k_mean(const double *__restrict__ spart, int nblocks, double invN,
                                                  double *__restrict__ out, const int *gate)
{
    if (gated_off_p(gate)) return;
    mean_partials<C>(spart, nblocks, invN, out);
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
template <typename R>
struct Launch {
    stz_ctx *ctx;
    SolveConsts k;
    int *nanflag;
    const SolveRep *reps = nullptr;   // batched: nrep replicas of ctx's shape
    int nrep = 1;
    hipStream_t stream = nullptr;

    // largest power-of-two row count <= rows_for(L0) that divides the row total
    int pick_rows(int L0max) const
    {
        int NR = L0max;
        while (NR > 2 && (k.rows % NR) != 0) NR /= 2;
        return (k.rows % NR) == 0 ? NR : 0;
    }

    // Each go_* launches a compiled variant or reports that none exists (never skips
    // silently).  Radix plans other than a length's default are compiled for the default
    // tile of the x and y passes, and for every tile of the fused pass.
    template <int L0, int NRV, int C, int PL>
    int go_fwd_x(dim3 grid)
    {
        if constexpr (x_variant<R, C>(L0, NRV) && plan_ok(L0, PL) && (PL == 0 || NRV == rows_for<R, C>(L0))) {
            hipLaunchKernelGGL((k_fwd_x<R, L0, NRV, C, 0, PL>), grid, dim3(NT_SOLVE), 0, stream, ctx->eps0,
                               (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[0], k, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int L0, int NRV, int C, int PL>
    int go_inv_x(dim3 grid, double *spart)
    {
        if constexpr (x_variant<R, C>(L0, NRV) && plan_ok(L0, PL) && (PL == 0 || NRV == rows_for<R, C>(L0))) {
            hipLaunchKernelGGL((k_inv_x<R, L0, NRV, C, 0, PL>), grid, dim3(NT_SOLVE), 0, stream,
                               (cpx<R> *)ctx->spec, ctx->eps0, (R *)ctx->sigma, spart, nanflag,
                               (const cpx<R> *)ctx->tw.tw[0], k, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    // the tuned row count when it is a compiled variant dividing the rows, else the default
    template <int L0, int C>
    int x_rows(int tuned) const
    {
        if (tuned > 0 && k.rows % tuned == 0 && x_variant<R, C>(L0, tuned)) return tuned;
        return pick_rows(rows_for<R, C>(L0));
    }
    // the tuned radix plan when it is compiled for this tile, else the default plan
    template <int L0, int C>
    static int x_plan(int tuned, int NR)
    {
        const int pl = tune_plan(tuned);
        return (pl > 0 && plan_ok(L0, pl) && NR == rows_for<R, C>(L0)) ? pl : 0;
    }

    template <int L0, int NRV, int C>
    int fwd_x_p(dim3 grid, int pl)
    {
        switch (pl) {
        case 0: return go_fwd_x<L0, NRV, C, 0>(grid);
        case 1: return go_fwd_x<L0, NRV, C, 1>(grid);
        case 2: return go_fwd_x<L0, NRV, C, 2>(grid);
        case 3: return go_fwd_x<L0, NRV, C, 3>(grid);
        default: return STZ_E_UNSUPPORTED;
        }
    }

    template <int L0, int NRV, int C>
    int inv_x_p(dim3 grid, double *spart, int pl)
    {
        switch (pl) {
        case 0: return go_inv_x<L0, NRV, C, 0>(grid, spart);
        case 1: return go_inv_x<L0, NRV, C, 1>(grid, spart);
        case 2: return go_inv_x<L0, NRV, C, 2>(grid, spart);
        case 3: return go_inv_x<L0, NRV, C, 3>(grid, spart);
        default: return STZ_E_UNSUPPORTED;
        }
    }

    template <int L0, int C>
    int fwd_x_n()
    {
        const int NR = x_rows<L0, C>(tune_tile(ctx->tune[0]));
        if (!NR) return STZ_E_ARG;
        const int pl = x_plan<L0, C>(ctx->tune[0], NR);
        const dim3 grid(k.rows / NR, nrep);
        int rc;
        switch (NR) {
        case 32: rc = fwd_x_p<L0, 32, C>(grid, pl); break;
        case 16: rc = fwd_x_p<L0, 16, C>(grid, pl); break;
        case 8: rc = fwd_x_p<L0, 8, C>(grid, pl); break;
        case 4: rc = fwd_x_p<L0, 4, C>(grid, pl); break;
        case 2: rc = fwd_x_p<L0, 2, C>(grid, pl); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled x-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }

    template <int L0, int C>
    int inv_x_n(double *spart, int *nblocks)
    {
        const int NR = x_rows<L0, C>(tune_tile(ctx->tune[3]));
        if (!NR) return STZ_E_ARG;
        const int pl = x_plan<L0, C>(ctx->tune[3], NR);
        *nblocks = k.rows / NR;
        const dim3 grid(k.rows / NR, nrep);
        int rc;
        switch (NR) {
        case 32: rc = inv_x_p<L0, 32, C>(grid, spart, pl); break;
        case 16: rc = inv_x_p<L0, 16, C>(grid, spart, pl); break;
        case 8: rc = inv_x_p<L0, 8, C>(grid, spart, pl); break;
        case 4: rc = inv_x_p<L0, 4, C>(grid, spart, pl); break;
        case 2: rc = inv_x_p<L0, 2, C>(grid, spart, pl); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled x-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }

    template <int L1, int W, int PL>
    int go_ypass(R sgn)
    {
        if constexpr (y_variant<R>(L1, W) && plan_ok(L1, PL) && (PL == 0 || W == cols_for(L1))) {
            const dim3 grid((k.h0 + W - 1) / W, k.n[2], ctx->C * nrep);
            hipLaunchKernelGGL((k_ypass<R, L1, W, PL>), grid, dim3(NT_SOLVE), 0, stream, (cpx<R> *)ctx->spec,
                               (const cpx<R> *)ctx->tw.tw[1], k, sgn, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int L1, int W>
    int ypass_p(R sgn, int pl)
    {
        switch (pl) {
        case 0: return go_ypass<L1, W, 0>(sgn);
        case 1: return go_ypass<L1, W, 1>(sgn);
        case 2: return go_ypass<L1, W, 2>(sgn);
        case 3: return go_ypass<L1, W, 3>(sgn);
        default: return STZ_E_UNSUPPORTED;
        }
    }

    template <int L1>
    int ypass_n(R sgn)
    {
        int W = tune_tile(ctx->tune[1]);
        if (!(W > 0 && y_variant<R>(L1, W))) W = cols_for(L1);
        int pl = tune_plan(ctx->tune[1]);
        if (!(pl > 0 && plan_ok(L1, pl) && W == cols_for(L1))) pl = 0;
        int rc;
        switch (W) {
        case 32: rc = ypass_p<L1, 32>(sgn, pl); break;
        case 16: rc = ypass_p<L1, 16>(sgn, pl); break;
        case 8: rc = ypass_p<L1, 8>(sgn, pl); break;
        case 4: rc = ypass_p<L1, 4>(sgn, pl); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled y-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }

    template <int LZ, int C, int DIM, int W, int PL>
    int go_fused()
    {
        if constexpr (f_variant<R, C>(LZ, W) && plan_ok(LZ, PL)) {
            constexpr int NTF = fused_threads(LZ * C * W);
            const dim3 grid((k.h0 + W - 1) / W, DIM == 3 ? k.n[1] : 1, nrep);
            hipLaunchKernelGGL((k_fused<R, LZ, W, C, DIM, NTF, 0, PL>), grid, dim3(NTF), 0, stream,
                               (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[DIM - 1], k, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int LZ, int C, int DIM, int W>
    int fused_p(int pl)
    {
        switch (pl) {
        case 0: return go_fused<LZ, C, DIM, W, 0>();
        case 1: return go_fused<LZ, C, DIM, W, 1>();
        case 2: return go_fused<LZ, C, DIM, W, 2>();
        case 3: return go_fused<LZ, C, DIM, W, 3>();
        default: return STZ_E_UNSUPPORTED;
        }
    }

    template <int LZ, int C, int DIM>
    int fused_n()
    {
        int W = tune_tile(ctx->tune[2]);
        if (!(W > 0 && f_variant<R, C>(LZ, W))) W = fused_cols_for<R, C>(LZ);
        int pl = tune_plan(ctx->tune[2]);
        if (!plan_ok(LZ, pl)) pl = 0;
        int rc;
        switch (W) {
        case 16: rc = fused_p<LZ, C, DIM, 16>(pl); break;
        case 8: rc = fused_p<LZ, C, DIM, 8>(pl); break;
        case 4: rc = fused_p<LZ, C, DIM, 4>(pl); break;
        case 2: rc = fused_p<LZ, C, DIM, 2>(pl); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled fused-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }
};

template <typename R>
int solve_impl(stz_ctx *ctx, int gated, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts)
{
    Launch<R> Lr;
    Lr.ctx = ctx;
    Lr.reps = reps;
    Lr.nrep = nrep;
    Lr.stream = stream;
    SolveConsts &k = Lr.k;
    k = solve_consts(ctx);
    k.gate = gated ? reinterpret_cast<const int *>(ctx->res + RES_KIND) : nullptr;
    Lr.nanflag = reinterpret_cast<int *>(ctx->res + 15);
    int rc = STZ_OK;
    const int C = ctx->C;
    tmark(ctx, 0);
    if (C == 3) {
        STZ_LEN_DISPATCH(ctx->n[0], rc = Lr.template fwd_x_n<LL, 3>());
        if (rc) return rc;
        tmark(ctx, 1);
        tmark(ctx, 2);
        STZ_LEN_DISPATCH(ctx->n[1], rc = (Lr.template fused_n<LL, 3, 2>()));
        if (rc) return rc;
        tmark(ctx, 3);
        tmark(ctx, 4);
    } else {
        STZ_LEN_DISPATCH(ctx->n[0], rc = Lr.template fwd_x_n<LL, 6>());
        if (rc) return rc;
        tmark(ctx, 1);
        STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)-1));
        if (rc) return rc;
        tmark(ctx, 2);
        STZ_LEN_DISPATCH(ctx->n[2], rc = (Lr.template fused_n<LL, 6, 3>()));
        if (rc) return rc;
        tmark(ctx, 3);
        STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)1));
        if (rc) return rc;
        tmark(ctx, 4);
    }
    int nblocks = 0;
    if (C == 3) {
        STZ_LEN_DISPATCH(ctx->n[0], rc = (Lr.template inv_x_n<LL, 3>(ctx->spart, &nblocks)));
    } else {
        STZ_LEN_DISPATCH(ctx->n[0], rc = (Lr.template inv_x_n<LL, 6>(ctx->spart, &nblocks)));
    }
    if (rc) return rc;
    tmark(ctx, 5);
    if (nparts) {           // batched: the caller reduces the partials
        *nparts = nblocks;
        return STZ_OK;
    }
    if (C == 3) hipLaunchKernelGGL((k_mean<3>), dim3(1), dim3(NT_MEAN), 0, stream, ctx->spart, nblocks, k.invN, ctx->res, k.gate);
    else hipLaunchKernelGGL((k_mean<6>), dim3(1), dim3(NT_MEAN), 0, stream, ctx->spart, nblocks, k.invN, ctx->res, k.gate);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 6);
    return STZ_OK;
}

// The spectral middle of a solve of ctx's own spectrum, ungated: the y passes (3-D) and
// the fused last-axis pass with Gamma-hat, with the context's tile choices.
template <typename R>
int spectral_impl(stz_ctx *ctx)
{
    Launch<R> Lr;
    Lr.ctx = ctx;
    Lr.stream = ctx->stream;
    Lr.k = solve_consts(ctx);
    Lr.nanflag = reinterpret_cast<int *>(ctx->res + 15);
    int rc = STZ_OK;
    if (ctx->C == 3) {
        STZ_LEN_DISPATCH(ctx->n[1], rc = (Lr.template fused_n<LL, 3, 2>()));
        return rc;
    }
    STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)-1));
    if (rc) return rc;
    STZ_LEN_DISPATCH(ctx->n[2], rc = (Lr.template fused_n<LL, 6, 3>()));
    if (rc) return rc;
    STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)1));
    return rc;
}

// Print-step displacement (HomogeneousStressSolver's displacement branch,
// STZ/elastic.c:873-900) of the context's eps0 (round-tripped) and EpsAvg into u[DIM][N]
// on the device: x-forward, y-forward, the fused z pass with the displacement operator,
// y-inverse, and the C2R x pass with the position term.  Default tiles (print time only).
// With per-voxel moduli the input is the eigenstrain the last inhomogeneous solve ended
// on, as the reference's final HomogeneousStressSolver(epsolast, uflag) call
// (STZ/elastic.c:1051-1055).
template <typename R>
int disp_impl(stz_ctx *ctx, double *u)
{
    Launch<R> Lr;
    Lr.ctx = ctx;
    Lr.stream = ctx->stream;
    SolveConsts &k = Lr.k;
    k = solve_consts(ctx);
    Lr.nanflag = reinterpret_cast<int *>(ctx->res + 15);
    const int saved[4] = {ctx->tune[0], ctx->tune[1], ctx->tune[2], ctx->tune[3]};
    int rc = STZ_OK;
    if (ctx->lamx) {
        // elastic inhomogeneity: the eigenstrain of the last fixed-point iteration
        // (the reference's epsolast) instead of the round-tripped eps0
        rc = stz_inhom_fwd_x(ctx, ctx->estar[ctx->inhom_cur]);
    } else {
        for (int t = 0; t < 4; ++t) ctx->tune[t] = 0;      // default tiles
        if (ctx->C == 3) {
            STZ_LEN_DISPATCH(ctx->n[0], rc = Lr.template fwd_x_n<LL, 3>());
        } else {
            STZ_LEN_DISPATCH(ctx->n[0], rc = Lr.template fwd_x_n<LL, 6>());
        }
        for (int t = 0; t < 4; ++t) ctx->tune[t] = saved[t];
    }
    if (rc) return rc;
    if (ctx->C == 6) {
        for (int t = 0; t < 4; ++t) ctx->tune[t] = 0;
        STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)-1));
        for (int t = 0; t < 4; ++t) ctx->tune[t] = saved[t];
        if (rc) return rc;
    }
    if (ctx->C == 3) {
        STZ_LEN_DISPATCH(ctx->n[1], {
            constexpr int W = fused_cols_for<R, 3>(LL);
            constexpr int NTF = fused_threads(LL * 3 * W);
            hipLaunchKernelGGL((k_fused<R, LL, W, 3, 2, NTF, 1>), dim3((k.h0 + W - 1) / W, 1, 1), dim3(NTF), 0,
                               ctx->stream, (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[1], k, nullptr);
        });
    } else {
        STZ_LEN_DISPATCH(ctx->n[2], {
            constexpr int W = fused_cols_for<R, 6>(LL);
            constexpr int NTF = fused_threads(LL * 6 * W);
            hipLaunchKernelGGL((k_fused<R, LL, W, 6, 3, NTF, 1>), dim3((k.h0 + W - 1) / W, k.n[1], 1), dim3(NTF), 0,
                               ctx->stream, (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[2], k, nullptr);
        });
    }
    STZ_CHECK_LAUNCH();
    if (ctx->C == 6) {
        for (int t = 0; t < 4; ++t) ctx->tune[t] = 0;
        STZ_LEN_DISPATCH(ctx->n[1], rc = Lr.template ypass_n<LL>((R)1));
        for (int t = 0; t < 4; ++t) ctx->tune[t] = saved[t];
        if (rc) return rc;
    }
    const int dim = ctx->dim;
    auto go = [&](auto L0c) -> int {
        constexpr int L0 = decltype(L0c)::value;
        const int NR = ctx->C == 3 ? Lr.pick_rows(rows_for<R, 3>(L0)) : Lr.pick_rows(rows_for<R, 6>(L0));
        if (!NR) return STZ_E_ARG;
        const dim3 grid(k.rows / NR);
        // NR is the default x-pass row count for C components (<= 32 KB of LDS), so the
        // DIM <= C component tile of each reachable variant fits
        auto launch = [&](auto NRc) -> int {
            constexpr int NRV = decltype(NRc)::value;
            if constexpr ((long long)L0 * (3 * NRV / 2 + 1) * (long long)sizeof(cpx<R>) <= 65536) {
                if (dim == 2)
                    hipLaunchKernelGGL((k_inv_x_disp<R, L0, NRV, 2>), grid, dim3(NT_SOLVE), 0, ctx->stream,
                                       (const cpx<R> *)ctx->spec, u, (const cpx<R> *)ctx->tw.tw[0], k);
                else
                    hipLaunchKernelGGL((k_inv_x_disp<R, L0, NRV, 3>), grid, dim3(NT_SOLVE), 0, ctx->stream,
                                       (const cpx<R> *)ctx->spec, u, (const cpx<R> *)ctx->tw.tw[0], k);
                return STZ_OK;
            } else {
                ctx->err = "no compiled displacement x-pass for this grid";
                return STZ_E_UNSUPPORTED;
            }
        };
        switch (NR) {
        case 32: return launch(std::integral_constant<int, 32>());
        case 16: return launch(std::integral_constant<int, 16>());
        case 8: return launch(std::integral_constant<int, 8>());
        case 4: return launch(std::integral_constant<int, 4>());
        case 2: return launch(std::integral_constant<int, 2>());
        default: return STZ_E_ARG;
        }
    };
    STZ_LEN_DISPATCH(ctx->n[0], rc = go(std::integral_constant<int, LL>()));
    if (rc) return rc;
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

}  // namespace

int solve_entry(STZ_SOLVE_REAL, stz_ctx *ctx, int gated, int nrep, const SolveRep *reps, hipStream_t stream,
                int *nparts)
{
    return solve_impl<STZ_SOLVE_REAL>(ctx, gated, nrep, reps, stream, nparts);
}

int spectral_entry(STZ_SOLVE_REAL, stz_ctx *ctx) { return spectral_impl<STZ_SOLVE_REAL>(ctx); }

int disp_entry(STZ_SOLVE_REAL, stz_ctx *ctx, double *u) { return disp_impl<STZ_SOLVE_REAL>(ctx, u); }
