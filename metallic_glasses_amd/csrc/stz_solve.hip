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
#include <vector>

namespace {

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

// The candidates of a slot: every compiled tile with the length's default radix plan,
// then the default tile (x, y passes) or every tile (fused pass) with each alternative
// plan (value tile | plan << 8).
template <typename R, int C>
int candidates_t(const stz_ctx *ctx, int slot, int *out)
{
    const int rows = ctx->n[1] * ctx->n[2];
    int n = 0;
    auto plans = [&](int L, int tile) {
        for (int j = 1; j < STZ_NPLAN; ++j)
            if (plan_ok(L, j)) out[n++] = tile | (j << 8);
    };
    if (slot == 0 || slot == 3) {
        const int L0 = ctx->n[0];
        for (int v = 32; v >= 2; v /= 2)
            if (x_variant<R, C>(L0, v) && rows % v == 0) out[n++] = v;
        if (rows % rows_for<R, C>(L0) == 0) plans(L0, rows_for<R, C>(L0));
    } else if (slot == 1) {
        if (ctx->dim != 3) return 0;
        for (int v = 32; v >= 4; v /= 2)
            if (y_variant<R>(ctx->n[1], v)) out[n++] = v;
        plans(ctx->n[1], cols_for(ctx->n[1]));
    } else {
        // every tile with every plan
        const int L = ctx->n[ctx->dim - 1];
        for (int v = 16; v >= 2; v /= 2)
            if (f_variant<R, C>(L, v)) out[n++] = v;
        for (int v = 16; v >= 2; v /= 2)
            if (f_variant<R, C>(L, v)) plans(L, v);
    }
    return n;
}

}  // namespace

// compiled tile variants of solve kernel slot (0 x forward, 1 y passes, 2 fused, 3 x inverse)
int tune_candidates(const stz_ctx *ctx, int slot, int out[STZ_TUNE_MAXCAND])
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
    return ctx->prec == STZ_FP64 ? solve_entry(0.0, ctx, gated, 1, nullptr, ctx->stream, nullptr)
                                 : solve_entry(0.0f, ctx, gated, 1, nullptr, ctx->stream, nullptr);
}

int stz_solve_spectral(stz_ctx *ctx)
{
    return ctx->prec == STZ_FP64 ? spectral_entry(0.0, ctx) : spectral_entry(0.0f, ctx);
}

int stz_displacement_launch(stz_ctx *ctx, double *u)
{
    return ctx->prec == STZ_FP64 ? disp_entry(0.0, ctx, u) : disp_entry(0.0f, ctx, u);
}

int stz_solve_launch_batched(stz_ctx *proto, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts)
{
    return proto->prec == STZ_FP64 ? solve_entry(0.0, proto, 1, nrep, reps, stream, nparts)
                                   : solve_entry(0.0f, proto, 1, nrep, reps, stream, nparts);
}

// rows of the stress-mean partial array: the largest block count of k_inv_x
long long stz_solve_partials(const stz_ctx *ctx)
{
    return (long long)ctx->n[1] * ctx->n[2];
}
