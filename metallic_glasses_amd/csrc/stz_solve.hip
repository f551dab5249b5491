// stz_solve.hip -- Green's-function stress solve on the GPU.
//
// Replaces InhomogeneousStressSolver(uflag=0) with uniform moduli
// (STZ/elastic.c:977-1066), i.e. one HomogeneousStressSolver (STZ/elastic.c:779-828)
// on the round-tripped eigenstrain C0^-1(C0:eps0) (STZ/elastic.c:999-1009).
//
// Data flow (one launch per box; C = 3 or 6 components):
//   k_fwd_x   eps0 (FP64) -> round trip -> R2C along x (two real rows packed per
//             complex line)                          -> spec[c][row][k0], k0 <= n0/2
//   k_ypass   complex FFT along y, in place (3-D only)
//   k_fused   last axis (z in 3-D, y in 2-D): forward FFT of all C components,
//             Gamma-hat multiply, inverse FFT -- the spectrum of one tile never
//             leaves the workgroup between the two transforms
//   k_ypass   inverse along y (3-D only)
//   k_inv_x   C2R along x + epilogue sigma = C0:(eps/N + epsavg - eps0_rt), block
//             partial sums of sigma (FP64)
//   k_mean    fixed-order reduction of the partials -> <sigma> (the resident step does
//             it in its bookkeeping kernel with the same code)
//
// Half-spectrum equivalence: the reference transforms the full complex spectrum and
// keeps Re(IFFT) (STZ/elastic.c:811-824).  Its k vectors use round() (STZ/init.c:
// 126-129), so on Nyquist planes Gamma-hat(n(k)) != Gamma-hat(n(-k)); Re() then applies
// the average of the two.  The half-spectrum path applies that average explicitly on
// Nyquist bins, which reproduces the reference to rounding.
#include "stz_solve_dev.h"

#include <cmath>
#include <type_traits>

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
bool supported_len(int L)
{
    switch (L) {
    case 4: case 8: case 16: case 32: case 64: case 96: case 128: case 256: case 512:
        return true;
    default:
        return false;
    }
}

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

    // each go_* launches a compiled variant or reports that none exists (never skips silently)
    template <int L0, int NRV, int C>
    int go_fwd_x(dim3 grid)
    {
        if constexpr (x_variant<R, C>(L0, NRV)) {
            hipLaunchKernelGGL((k_fwd_x<R, L0, NRV, C>), grid, dim3(NT_SOLVE), 0, stream, ctx->eps0,
                               (cpx<R> *)ctx->spec, (const cpx<R> *)ctx->tw.tw[0], k, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int L0, int NRV, int C>
    int go_inv_x(dim3 grid, double *spart)
    {
        if constexpr (x_variant<R, C>(L0, NRV)) {
            hipLaunchKernelGGL((k_inv_x<R, L0, NRV, C>), grid, dim3(NT_SOLVE), 0, stream, (cpx<R> *)ctx->spec,
                               ctx->eps0, (R *)ctx->sigma, spart, nanflag, (const cpx<R> *)ctx->tw.tw[0], k, reps);
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

    template <int L0, int C>
    int fwd_x_n()
    {
        const int NR = x_rows<L0, C>(ctx->tune[0]);
        if (!NR) return STZ_E_ARG;
        const dim3 grid(k.rows / NR, nrep);
        int rc;
        switch (NR) {
        case 32: rc = go_fwd_x<L0, 32, C>(grid); break;
        case 16: rc = go_fwd_x<L0, 16, C>(grid); break;
        case 8: rc = go_fwd_x<L0, 8, C>(grid); break;
        case 4: rc = go_fwd_x<L0, 4, C>(grid); break;
        case 2: rc = go_fwd_x<L0, 2, C>(grid); break;
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
        const int NR = x_rows<L0, C>(ctx->tune[3]);
        if (!NR) return STZ_E_ARG;
        *nblocks = k.rows / NR;
        const dim3 grid(k.rows / NR, nrep);
        int rc;
        switch (NR) {
        case 32: rc = go_inv_x<L0, 32, C>(grid, spart); break;
        case 16: rc = go_inv_x<L0, 16, C>(grid, spart); break;
        case 8: rc = go_inv_x<L0, 8, C>(grid, spart); break;
        case 4: rc = go_inv_x<L0, 4, C>(grid, spart); break;
        case 2: rc = go_inv_x<L0, 2, C>(grid, spart); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled x-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }

    template <int L1, int W>
    int go_ypass(R sgn)
    {
        if constexpr (y_variant<R>(L1, W)) {
            const dim3 grid((k.h0 + W - 1) / W, k.n[2], ctx->C * nrep);
            hipLaunchKernelGGL((k_ypass<R, L1, W>), grid, dim3(NT_SOLVE), 0, stream, (cpx<R> *)ctx->spec,
                               (const cpx<R> *)ctx->tw.tw[1], k, sgn, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int L1>
    int ypass_n(R sgn)
    {
        int W = ctx->tune[1];
        if (!(W > 0 && y_variant<R>(L1, W))) W = cols_for(L1);
        int rc;
        switch (W) {
        case 32: rc = go_ypass<L1, 32>(sgn); break;
        case 16: rc = go_ypass<L1, 16>(sgn); break;
        case 8: rc = go_ypass<L1, 8>(sgn); break;
        case 4: rc = go_ypass<L1, 4>(sgn); break;
        default: rc = STZ_E_UNSUPPORTED; break;
        }
        if (rc) {
            ctx->err = "no compiled y-pass variant for this grid";
            return rc;
        }
        STZ_CHECK_LAUNCH();
        return STZ_OK;
    }

    template <int LZ, int C, int DIM, int W>
    int go_fused()
    {
        if constexpr (f_variant<R, C>(LZ, W)) {
            constexpr int NTF = fused_threads(LZ * C * W);
            const dim3 grid((k.h0 + W - 1) / W, DIM == 3 ? k.n[1] : 1, nrep);
            hipLaunchKernelGGL((k_fused<R, LZ, W, C, DIM, NTF>), grid, dim3(NTF), 0, stream, (cpx<R> *)ctx->spec,
                               (const cpx<R> *)ctx->tw.tw[DIM - 1], k, reps);
            return STZ_OK;
        } else {
            return STZ_E_UNSUPPORTED;
        }
    }

    template <int LZ, int C, int DIM>
    int fused_n()
    {
        int W = ctx->tune[2];
        if (!(W > 0 && f_variant<R, C>(LZ, W))) W = fused_cols_for<R, C>(LZ);
        int rc;
        switch (W) {
        case 16: rc = go_fused<LZ, C, DIM, 16>(); break;
        case 8: rc = go_fused<LZ, C, DIM, 8>(); break;
        case 4: rc = go_fused<LZ, C, DIM, 4>(); break;
        case 2: rc = go_fused<LZ, C, DIM, 2>(); break;
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

template <typename R>
int upload_twiddles(stz_ctx *ctx, int axis, int L)
{
    std::vector<cpx<R>> h(L);
    for (int t = 0; t < L; ++t) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)L;
        h[t].x = (R)(double)cosl(a);
        h[t].y = (R)(-(double)sinl(a));
    }
    STZ_HIP(hipMalloc(&ctx->tw.tw[axis], sizeof(cpx<R>) * L));
    STZ_HIP(hipMemcpy(ctx->tw.tw[axis], h.data(), sizeof(cpx<R>) * L, hipMemcpyHostToDevice));
    return STZ_OK;
}

}  // namespace

namespace {

template <typename R, int C>
int candidates_t(const stz_ctx *ctx, int slot, int *out)
{
    const int rows = ctx->n[1] * ctx->n[2];
    int n = 0;
    if (slot == 0 || slot == 3) {
        const int L0 = ctx->n[0];
        for (int v = 32; v >= 2; v /= 2)
            if (x_variant<R, C>(L0, v) && rows % v == 0) out[n++] = v;
    } else if (slot == 1) {
        if (ctx->dim != 3) return 0;
        for (int v = 32; v >= 4; v /= 2)
            if (y_variant<R>(ctx->n[1], v)) out[n++] = v;
    } else {
        const int L = ctx->n[ctx->dim - 1];
        for (int v = 16; v >= 2; v /= 2)
            if (f_variant<R, C>(L, v)) out[n++] = v;
    }
    return n;
}

}  // namespace

// compiled tile variants of solve kernel slot (0 x forward, 1 y passes, 2 fused, 3 x inverse)
int tune_candidates(const stz_ctx *ctx, int slot, int out[8])
{
    if (ctx->prec == STZ_FP64)
        return ctx->C == 3 ? candidates_t<double, 3>(ctx, slot, out) : candidates_t<double, 6>(ctx, slot, out);
    return ctx->C == 3 ? candidates_t<float, 3>(ctx, slot, out) : candidates_t<float, 6>(ctx, slot, out);
}

bool stz_solve_supported(int dim, const int *n, int prec)
{
    (void)prec;
    for (int a = 0; a < dim; ++a)
        if (!supported_len(n[a])) return false;
    if (dim == 2 && n[2] != 1) return false;
    return true;
}

int stz_fft_setup(stz_ctx *ctx)
{
    for (int a = 0; a < ctx->dim; ++a) {
        int rc = ctx->prec == STZ_FP64 ? upload_twiddles<double>(ctx, a, ctx->n[a])
                                       : upload_twiddles<float>(ctx, a, ctx->n[a]);
        if (rc) return rc;
    }
    return STZ_OK;
}

void stz_fft_free(stz_ctx *ctx)
{
    for (int a = 0; a < 3; ++a) {
        if (ctx->tw.tw[a]) (void)hipFree(ctx->tw.tw[a]);
        ctx->tw.tw[a] = nullptr;
    }
}

int stz_solve_launch(stz_ctx *ctx, int gated)
{
    return ctx->prec == STZ_FP64 ? solve_impl<double>(ctx, gated, 1, nullptr, ctx->stream, nullptr)
                                 : solve_impl<float>(ctx, gated, 1, nullptr, ctx->stream, nullptr);
}

int stz_solve_spectral(stz_ctx *ctx)
{
    return ctx->prec == STZ_FP64 ? spectral_impl<double>(ctx) : spectral_impl<float>(ctx);
}

int stz_displacement_launch(stz_ctx *ctx, double *u)
{
    return ctx->prec == STZ_FP64 ? disp_impl<double>(ctx, u) : disp_impl<float>(ctx, u);
}

int stz_solve_launch_batched(stz_ctx *proto, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts)
{
    return proto->prec == STZ_FP64 ? solve_impl<double>(proto, 1, nrep, reps, stream, nparts)
                                   : solve_impl<float>(proto, 1, nrep, reps, stream, nparts);
}

// rows of the stress-mean partial array: the largest block count of k_inv_x
long long stz_solve_partials(const stz_ctx *ctx)
{
    return (long long)ctx->n[1] * ctx->n[2];
}
