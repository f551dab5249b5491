// stz_solve_dev.h -- device pieces shared by the stress-solve translation units: the
// launch constants, the per-replica view of a batched launch, the elementwise elasticity
// of the reference (literal expression order), and the two x-axis passes (R2C + round
// trip, C2R + stress epilogue) in their uniform (MODE 0) and inhomogeneous (MODE 1) forms.
#pragma once
#include "stz_internal.h"

// compile-time length of a runtime axis length (the supported FFT lengths)
#define STZ_LEN_DISPATCH(L, ...)                                \
    switch (L) {                                                \
    case 4: { constexpr int LL = 4; __VA_ARGS__; } break;       \
    case 8: { constexpr int LL = 8; __VA_ARGS__; } break;       \
    case 16: { constexpr int LL = 16; __VA_ARGS__; } break;     \
    case 32: { constexpr int LL = 32; __VA_ARGS__; } break;     \
    case 64: { constexpr int LL = 64; __VA_ARGS__; } break;     \
    case 96: { constexpr int LL = 96; __VA_ARGS__; } break;     \
    case 128: { constexpr int LL = 128; __VA_ARGS__; } break;   \
    case 256: { constexpr int LL = 256; __VA_ARGS__; } break;   \
    case 512: { constexpr int LL = 512; __VA_ARGS__; } break;   \
    default: return STZ_E_ARG;                                  \
    }

namespace {

constexpr int NT_SOLVE = 256;

struct SolveConsts {
    const int *gate;    // fused step: run only when the step kind is thermal
    double lambda, mu, alpha, s0, sh, sa, invN;
    double epsavg[6];
    int n[3];
    double rn[3];       // 1.0 / n[a] (the reference grid's k-vector scale, computed once)
    long long N;
    int rows;           // n1*n2
    int h0, hp;
    int C;
    double L[3];        // voxel edges (displacement position term)
    // fixed-point iteration of the inhomogeneous solve (x passes of MODE 1 only)
    const double *tau0; // C(x):eps0 [C][N]
    const double *lamx, *mux;   // Lambda(x), Mu(x) [N]
    double *eout;       // next reference-medium eigenstrain [C][N]
    double *cpart;      // per-block sums of its squared change
    double hdt;         // heat conduction: exp(-|k| hdt) multiplier (k_fused OP 2)
};

// Per-replica view of a batched launch: the replica index is a grid coordinate, and
// its fields, uniform strain and step kind come from reps[r].  With reps == nullptr the
// launch is the single context's own (pointers from the arguments).
template <typename R>
struct View {
    const double *eps0;
    cpx<R> *spec;
    R *sigma;
    double *spart;
    int *nanflag;
    const int *gate;
    double ea[6];
};

template <typename R>
__device__ __forceinline__ View<R> view_of(const SolveRep *reps, int r, const SolveConsts &k, const double *eps0,
                                           cpx<R> *spec, R *sigma, double *spart, int *nanflag)
{
    View<R> v;
    if (reps) {
        const SolveRep &q = reps[r];
        v.eps0 = q.eps0;
        v.spec = (cpx<R> *)q.spec;
        v.sigma = (R *)q.sigma;
        v.spart = q.spart;
        v.nanflag = reinterpret_cast<int *>(q.res + 15);
        v.gate = reinterpret_cast<const int *>(q.res + RES_FFT);   // KIND_NONE when the step updates locally
#pragma unroll
        for (int c = 0; c < 6; ++c) v.ea[c] = q.epsavg[c];
    } else {
        v.eps0 = eps0;
        v.spec = spec;
        v.sigma = sigma;
        v.spart = spart;
        v.nanflag = nanflag;
        v.gate = k.gate;
#pragma unroll
        for (int c = 0; c < 6; ++c) v.ea[c] = k.epsavg[c];
    }
    return v;
}

// --- elementwise elasticity (literal expression order of the reference) ----------
#pragma clang fp contract(off)

// Tau0 = C0:e (IsotropicMul_RF with uniform moduli, STZ/elastic.c:905-920), then
// C0^-1(Tau0 + Tau1) with Tau1 = (C0 - C0):(e + epsavg) (IsotropicMul_RF_D :922-938,
// IsotropicInv :940-955).
template <int C>
__device__ __forceinline__ void round_trip(const double *e, const double *ea, double lam, double mu, double *r)
{
    double t0[C], t1[C], tau[C];
    double z[C];
#pragma unroll
    for (int c = 0; c < C; ++c) z[c] = e[c] + ea[c];
    if constexpr (C == 3) {
        double tra = lam * (e[0] + e[1]);
        t0[0] = tra + 2 * mu * e[0];
        t0[1] = tra + 2 * mu * e[1];
        t0[2] = 2 * mu * e[2];
        double tra1 = (lam - lam) * (z[0] + z[1]);
        t1[0] = tra1 + 2 * (mu - mu) * z[0];
        t1[1] = tra1 + 2 * (mu - mu) * z[1];
        t1[2] = 2 * (mu - mu) * z[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) tau[c] = t0[c] + t1[c];
        double tr = -1.0 * lam / 2.0 / mu / (2.0 * lam + 2.0 * mu) * (tau[0] + tau[1]);
        r[0] = tr + tau[0] / 2.0 / mu;
        r[1] = tr + tau[1] / 2.0 / mu;
        r[2] = tau[2] / 2.0 / mu;
    } else {
        double tra = lam * ((e[0] + e[1]) + e[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) t0[c] = tra + 2 * mu * e[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) t0[c] = 2 * mu * e[c];
        double tra1 = (lam - lam) * ((z[0] + z[1]) + z[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) t1[c] = tra1 + 2 * (mu - mu) * z[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) t1[c] = 2 * (mu - mu) * z[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) tau[c] = t0[c] + t1[c];
        double tr = -1.0 * lam / 2.0 / mu / (3.0 * lam + 2.0 * mu) * ((tau[0] + tau[1]) + tau[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = tr + tau[c] / 2.0 / mu;
#pragma unroll
        for (int c = 3; c < 6; ++c) r[c] = tau[c] / 2.0 / mu;
    }
}

// IsotropicMul_R2 (STZ/elastic.c:686-701): sigma = C0:d
template <int C>
__device__ __forceinline__ void stress_of(const double *d, double lam, double mu, double *s)
{
    if constexpr (C == 3) {
        double tra = lam * (d[0] + d[1]);
        s[0] = tra + 2.0 * mu * d[0];
        s[1] = tra + 2.0 * mu * d[1];
        s[2] = 2.0 * mu * d[2];
    } else {
        double tra = lam * ((d[0] + d[1]) + d[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] = tra + 2.0 * mu * d[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) s[c] = 2.0 * mu * d[c];
    }
}

// Gamma-hat applied to one real part (re or im) of the eigenstrain spectrum:
// IsotropicMul_C, HadamardProdnSum_C x3, HadamardProdnSSum_C x2, HadamardProd2nSSum_C
// (STZ/elastic.c:802-809); g = unit k vector, gg its products.
template <typename R, int C>
__device__ __forceinline__ void gamma_part(const R *e, const R *g, const R *gg, R lam, R mu, R s0, R sh, R sa, R *o)
{
    R t[C];
    if constexpr (C == 3) {
        R tra = lam * (e[0] + e[1]);
        t[0] = tra + 2 * mu * e[0];
        t[1] = tra + 2 * mu * e[1];
        t[2] = 2 * mu * e[2];
        R F0 = t[0] * g[0] + t[2] * g[1];
        R F1 = t[2] * g[0] + t[1] * g[1];
        R G = F0 * g[0] + F1 * g[1];
        o[0] = s0 * (F0 * g[0]) + sa * (G * gg[0]);
        o[1] = s0 * (F1 * g[1]) + sa * (G * gg[1]);
        o[2] = sh * (F0 * g[1] + F1 * g[0]) + sa * (G * gg[2]);
    } else {
        R tra = lam * ((e[0] + e[1]) + e[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = tra + 2 * mu * e[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) t[c] = 2 * mu * e[c];
        // Voigt (xx, yy, zz, yz, xz, xy)
        R F0 = t[0] * g[0] + t[5] * g[1] + t[4] * g[2];
        R F1 = t[5] * g[0] + t[1] * g[1] + t[3] * g[2];
        R F2 = t[4] * g[0] + t[3] * g[1] + t[2] * g[2];
        R G = F0 * g[0] + F1 * g[1] + F2 * g[2];
        o[0] = s0 * (F0 * g[0]) + sa * (G * gg[0]);
        o[1] = s0 * (F1 * g[1]) + sa * (G * gg[1]);
        o[2] = s0 * (F2 * g[2]) + sa * (G * gg[2]);
        o[3] = sh * (F1 * g[2] + F2 * g[1]) + sa * (G * gg[3]);
        o[4] = sh * (F0 * g[2] + F2 * g[0]) + sa * (G * gg[4]);
        o[5] = sh * (F0 * g[1] + F1 * g[0]) + sa * (G * gg[5]);
    }
}

// unit vector of the reference k grid (Grad, STZ/init.c:116-137) for integer indices;
// returns |k| (g2, set to 1 at k = 0); rn[a] = 1.0 / n[a]
template <int DIM>
__device__ __forceinline__ double unit_k(const int *idx, const double *rn, double *g)
{
    g[0] = g[1] = g[2] = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        double s = (double)(idx[a]) * rn[a];
        g[a] = (s - round(s)) * (2 * STZ_PI_REF);
    }
    double g2;
    if constexpr (DIM == 2) g2 = sqrt(g[0] * g[0] + g[1] * g[1]);
    else g2 = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (idx[0] + idx[1] + (DIM == 3 ? idx[2] : 0) == 0) g2 = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] /= g2;
    return g2;
}

// Displacement branch of HomogeneousStressSolver (uflag 1/2, STZ/elastic.c:873-887) on one
// real part of the eigenstrain spectrum: F = (C0:e).n, G = n.F and
// tmp_a = F_a + (-alpha)*(G*n_a) (HadamardSumnSProd, :765-777), a < DIM
template <typename R, int C, int DIM>
__device__ __forceinline__ void disp_part(const R *e, const R *g, R lam, R mu, R nalpha, R *o)
{
    R t[C], F[3], G;
    if constexpr (C == 3) {
        R tra = lam * (e[0] + e[1]);
        t[0] = tra + 2 * mu * e[0];
        t[1] = tra + 2 * mu * e[1];
        t[2] = 2 * mu * e[2];
        F[0] = t[0] * g[0] + t[2] * g[1];
        F[1] = t[2] * g[0] + t[1] * g[1];
        G = F[0] * g[0] + F[1] * g[1];
    } else {
        R tra = lam * ((e[0] + e[1]) + e[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = tra + 2 * mu * e[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) t[c] = 2 * mu * e[c];
        F[0] = t[0] * g[0] + t[5] * g[1] + t[4] * g[2];
        F[1] = t[5] * g[0] + t[1] * g[1] + t[3] * g[2];
        F[2] = t[4] * g[0] + t[3] * g[1] + t[2] * g[2];
        G = F[0] * g[0] + F[1] * g[1] + F[2] * g[2];
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) o[a] = F[a] + nalpha * (G * g[a]);
}

// supercell position added to the displacement (STZ/elastic.c:893-899).  2-D as the
// reference writes it: its slow index i along x, the fast index j along y; 3-D: x_b the
// grid index along axis b (x = the fast axis), u_a += sum_b x_b (delta_ab + e_ab) L_b.
template <int DIM>
__device__ __forceinline__ double position(int a, int i0, int i1, int i2, const double *e, const double *L)
{
    if constexpr (DIM == 2) {
        const double i = (double)i1, j = (double)i0;
        return a == 0 ? i * (1 + e[0]) * L[0] + j * e[2] * L[1] : i * e[2] * L[0] + j * (1 + e[1]) * L[1];
    } else {
        constexpr int vo[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
        const double x[3] = {(double)i0, (double)i1, (double)i2};
        double pos = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) pos = pos + x[b] * ((a == b ? 1.0 : 0.0) + e[vo[a][b]]) * L[b];
        return pos;
    }
}

template <typename R, int DIM>
__device__ __forceinline__ void products(const R *g, R *gg)
{
    if constexpr (DIM == 2) {
        gg[0] = g[0] * g[0];
        gg[1] = g[1] * g[1];
        gg[2] = g[0] * g[1];
    } else {
        gg[0] = g[0] * g[0];
        gg[1] = g[1] * g[1];
        gg[2] = g[2] * g[2];
        gg[3] = g[1] * g[2];
        gg[4] = g[0] * g[2];
        gg[5] = g[0] * g[1];
    }
}

// C:e for isotropic moduli (lam, mu) (IsotropicMul_RF / IsotropicMul_RF_D with the
// per-voxel moduli or their differences, STZ/elastic.c:905-938)
template <int C>
__device__ __forceinline__ void iso_mul(double lam, double mu, const double *e, double *t)
{
    if constexpr (C == 3) {
        double tra = lam * (e[0] + e[1]);
        t[0] = tra + 2 * mu * e[0];
        t[1] = tra + 2 * mu * e[1];
        t[2] = 2 * mu * e[2];
    } else {
        double tra = lam * ((e[0] + e[1]) + e[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = tra + 2 * mu * e[c];
#pragma unroll
        for (int c = 3; c < 6; ++c) t[c] = 2 * mu * e[c];
    }
}

// C0^-1 : t (IsotropicInv, STZ/elastic.c:940-955)
template <int C>
__device__ __forceinline__ void iso_inv(double lam, double mu, const double *t, double *e)
{
    if constexpr (C == 3) {
        double tra = -1.0 * lam / 2.0 / mu / (2.0 * lam + 2.0 * mu) * (t[0] + t[1]);
        e[0] = tra + t[0] / 2.0 / mu;
        e[1] = tra + t[1] / 2.0 / mu;
        e[2] = t[2] / 2.0 / mu;
    } else {
        double tra = -1.0 * lam / 2.0 / mu / (3.0 * lam + 2.0 * mu) * ((t[0] + t[1]) + t[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c] = tra + t[c] / 2.0 / mu;
#pragma unroll
        for (int c = 3; c < 6; ++c) e[c] = t[c] / 2.0 / mu;
    }
}

// One fixed-point update of the inhomogeneous solve at a voxel (STZ/elastic.c:1027-1046):
// from the total strain `tot` of the homogeneous solve of the eigenstrain `ein`,
// new = C0^-1 (Tau0 + (C0 - C(x)):tot), mixed 0.5/0.5 with `ein`.
template <int C>
__device__ __forceinline__ void inhom_update(const double *tot, const double *ein, const double *tau0, double lx,
                                             double mx, double lam, double mu, double *eout)
{
    double t1[C], tau[C], r[C];
    iso_mul<C>(lam - lx, mu - mx, tot, t1);
#pragma unroll
    for (int c = 0; c < C; ++c) tau[c] = tau0[c] + t1[c];
    iso_inv<C>(lam, mu, tau, r);
    const double mix = 0.5;
#pragma unroll
    for (int c = 0; c < C; ++c) eout[c] = mix * r[c] + (1 - mix) * ein[c];
}

#pragma clang fp contract(on)

// fused step: the solve runs after a thermal event or an athermal avalanche
__device__ __forceinline__ bool gated_off_p(const int *gate)
{
    return gate && *gate != KIND_THERMAL && *gate != KIND_ATHERMAL;
}

// ---------------------------------------------------------------------------------
// k_fwd_x: eps0 -> round trip -> R2C along x for NR rows and all C components
// (MODE 1: the input is the reference-medium eigenstrain itself, no round trip)
// ---------------------------------------------------------------------------------
template <typename R, int L0, int NR, int C, int MODE = 0, int PL = 0>
__global__ __launch_bounds__(NT_SOLVE) void k_fwd_x(const double *__restrict__ eps0_a, cpx<R> *__restrict__ spec_a,
                                                    const cpx<R> *__restrict__ twg, SolveConsts k, const SolveRep *reps)
{
    constexpr int NP = NR / 2;            // packed complex lines per component
    const View<R> vw = view_of<R>(reps, blockIdx.y, k, eps0_a, spec_a, nullptr, nullptr, nullptr);
    if (gated_off_p(vw.gate)) return;
    const double *__restrict__ eps0 = vw.eps0;
    cpx<R> *__restrict__ spec = vw.spec;
    constexpr int NL = C * NP;
    constexpr int LS = NL + 1;            // odd stride against bank conflicts
    __shared__ cpx<R> buf[L0 * LS];
    __shared__ cpx<R> tw[L0];
    const int tid = threadIdx.x;
    for (int i = tid; i < L0; i += NT_SOLVE) tw[i] = twg[i];
    const long long row0 = (long long)blockIdx.x * NR;
    R *bufr = reinterpret_cast<R *>(buf);
    struct Ev {
        double v[C];
    };
    staged_loop<Ev, NR * L0, NT_SOLVE, 4>(
        [&](int e) {
            const long long v = (row0 + e / L0) * L0 + e % L0;
            Ev r;
#pragma unroll
            for (int c = 0; c < C; ++c) r.v[c] = gload(eps0 + (long long)c * k.N + v);
            return r;
        },
        [&](int e, const Ev &ev) {
            const int rl = e / L0, i0 = e % L0;
            double rt[C];
            if constexpr (MODE == 0) round_trip<C>(ev.v, vw.ea, k.lambda, k.mu, rt);
            else
#pragma unroll
                for (int c = 0; c < C; ++c) rt[c] = ev.v[c];
#pragma unroll
            for (int c = 0; c < C; ++c) bufr[2 * (i0 * LS + c * NP + (rl >> 1)) + (rl & 1)] = (R)rt[c];
        });
    __syncthreads();
    fft_lines<R, L0, NL, LS, NT_SOLVE, PL>(buf, tw, (R)-1);
    constexpr int H0 = L0 / 2 + 1;
    for (int e = tid; e < C * NR * H0; e += NT_SOLVE) {
        const int kk = e % H0;
        const int rl = (e / H0) % NR;
        const int c = e / (H0 * NR);
        const int p = rl >> 1;
        const cpx<R> Z = buf[kk * LS + c * NP + p];
        const cpx<R> Zm = buf[((L0 - kk) % L0) * LS + c * NP + p];
        cpx<R> X;
        if ((rl & 1) == 0) X = {(Z.x + Zm.x) * (R)0.5, (Z.y - Zm.y) * (R)0.5};
        else X = {(Z.y + Zm.y) * (R)0.5, -(Z.x - Zm.x) * (R)0.5};
        gstore(spec + ((long long)c * k.rows + row0 + rl) * k.hp + kk, X);
    }
}

// ---------------------------------------------------------------------------------
// k_inv_x: C2R along x + stress epilogue + per-block partial sums of sigma
// (MODE 1: the input is the reference-medium eigenstrain; the epilogue also takes the
// fixed-point update of the inhomogeneous solve into k.eout and sums its squared change)
// ---------------------------------------------------------------------------------
template <typename R, int L0, int NR, int C, int MODE = 0, int PL = 0>
__global__ __launch_bounds__(NT_SOLVE) void k_inv_x(cpx<R> *__restrict__ spec_a, const double *__restrict__ eps0_a,
                                                    R *__restrict__ sigma_a, double *__restrict__ spart_a,
                                                    int *__restrict__ nanflag_a, const cpx<R> *__restrict__ twg,
                                                    SolveConsts k, const SolveRep *reps)
{
    constexpr int NP = NR / 2;
    constexpr int NL = C * NP;
    constexpr int LS = NL + 1;
    constexpr int H = L0 / 2;
    const View<R> vw = view_of<R>(reps, blockIdx.y, k, eps0_a, spec_a, sigma_a, spart_a, nanflag_a);
    if (gated_off_p(vw.gate)) return;
    const cpx<R> *__restrict__ spec = vw.spec;
    const double *__restrict__ eps0 = vw.eps0;
    R *__restrict__ sigma = vw.sigma;
    double *__restrict__ spart = vw.spart;
    int *__restrict__ nanflag = vw.nanflag;
    // the block reduction reuses the FFT tile when it is large enough
    constexpr int NRED = C + (MODE == 1 ? 1 : 0);
    constexpr bool RED_IN_BUF = (long long)L0 * LS * (long long)sizeof(cpx<R>) >= (long long)NRED * NT_SOLVE * 8;
    __shared__ cpx<R> buf[L0 * LS];
    __shared__ cpx<R> tw[L0];
    __shared__ double redx[RED_IN_BUF ? 1 : NRED * NT_SOLVE];
    const int tid = threadIdx.x;
    for (int i = tid; i < L0; i += NT_SOLVE) tw[i] = twg[i];
    const long long row0 = (long long)blockIdx.x * NR;
    struct Pair {
        cpx<R> a, b;
    };
    staged_loop<Pair, C * NP * L0, NT_SOLVE, 6>(
        [&](int e) {
            const int kk = e % L0;
            const int p = (e / L0) % NP;
            const int c = e / (L0 * NP);
            const cpx<R> *ra = spec + ((long long)c * k.rows + row0 + 2 * p) * k.hp + (kk <= H ? kk : L0 - kk);
            Pair q;
            q.a = gload(ra);
            q.b = gload(ra + k.hp);
            return q;
        },
        [&](int e, const Pair &q) {
            const int kk = e % L0;
            const int p = (e / L0) % NP;
            const int c = e / (L0 * NP);
            cpx<R> Z;
            if (kk == 0 || kk == H) Z = {q.a.x, q.b.x};
            else if (kk < H) Z = {q.a.x - q.b.y, q.a.y + q.b.x};
            else Z = {q.a.x + q.b.y, -q.a.y + q.b.x};
            buf[kk * LS + c * NP + p] = Z;
        });
    __syncthreads();
    fft_lines<R, L0, NL, LS, NT_SOLVE, PL>(buf, tw, (R)1);
    const R *bufr = reinterpret_cast<const R *>(buf);
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    double cacc = 0.0;
    int bad = 0;
    struct In {
        double ev[C];
        double t0[MODE == 1 ? C : 1];
        double lx, mx;
    };
    staged_loop<In, NR * L0, NT_SOLVE, 2>(
        [&](int e) {
            const long long v = (row0 + e / L0) * L0 + e % L0;
            In r;
#pragma unroll
            for (int c = 0; c < C; ++c) r.ev[c] = gload(eps0 + (long long)c * k.N + v);
            if constexpr (MODE == 1) {
#pragma unroll
                for (int c = 0; c < C; ++c) r.t0[c] = gload(k.tau0 + (long long)c * k.N + v);
                r.lx = gload(k.lamx + v);
                r.mx = gload(k.mux + v);
            }
            return r;
        },
        [&](int e, const In &in) {
            const int rl = e / L0, i0 = e % L0;
            const long long v = (row0 + rl) * L0 + i0;
            double rt[C], d[C], s[C], tot[C];
            if constexpr (MODE == 0) round_trip<C>(in.ev, vw.ea, k.lambda, k.mu, rt);
            else
#pragma unroll
                for (int c = 0; c < C; ++c) rt[c] = in.ev[c];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const double z = (double)bufr[2 * (i0 * LS + c * NP + (rl >> 1)) + (rl & 1)];
                bad |= (z != z);
                tot[c] = z / (double)k.N + vw.ea[c];   // STZ/elastic.c:822
                d[c] = tot[c] - rt[c];
            }
            stress_of<C>(d, k.lambda, k.mu, s);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                gstore(sigma + (long long)c * k.N + v, (R)s[c]);
                acc[c] += s[c];
            }
            if constexpr (MODE == 1) {
                double nx[C];
                inhom_update<C>(tot, rt, in.t0, in.lx, in.mx, k.lambda, k.mu, nx);
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    gstore(k.eout + (long long)c * k.N + v, nx[c]);
                    const double dd = nx[c] - rt[c];
                    cacc += dd * dd;
                }
            }
        });
    if (bad) atomicOr(nanflag, 1);
    // fixed-shape tree reduction, all components at once (each component's tree is the
    // same as a reduction of that component alone)
    __syncthreads();   // every epilogue read of the tile is done before it is reused
    double *red = RED_IN_BUF ? reinterpret_cast<double *>(buf) : redx;
#pragma unroll
    for (int c = 0; c < C; ++c) red[c * NT_SOLVE + tid] = acc[c];
    if constexpr (MODE == 1) red[C * NT_SOLVE + tid] = cacc;
    __syncthreads();
    for (int w = NT_SOLVE / 2; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int c = 0; c < NRED; ++c) red[c * NT_SOLVE + tid] = red[c * NT_SOLVE + tid] + red[c * NT_SOLVE + tid + w];
        __syncthreads();
    }
    if (tid < C) spart[(long long)blockIdx.x * C + tid] = red[tid * NT_SOLVE];
    if constexpr (MODE == 1)
        if (tid == 0) k.cpart[blockIdx.x] = red[C * NT_SOLVE];
}

// rows per workgroup of the x passes: the largest power of two (2..32) whose LDS tile,
// L0 x (C*NR/2 + 1) complex values, stays within 32 KB
template <typename R, int C>
constexpr int rows_for(int L0)
{
    int nr = 32;
    while (nr > 2 && (long long)L0 * (C * nr / 2 + 1) * (long long)sizeof(cpx<R>) > 32768) nr /= 2;
    return nr;
}
template <typename R, int C>
constexpr bool x_variant(int L0, int NR)
{
    return NR >= 2 && NR <= 2 * rows_for<R, C>(L0) &&
           (long long)L0 * (C * NR / 2 + 1) * (long long)sizeof(cpx<R>) <= 65536;
}

// The workgroups of a launch are dispatched to the chip's 8 XCDs in turn (linear id
// mod 8), and each XCD has its own L2.  The column tiles of the strided passes are
// narrower than a 128-byte line (W complex values of a row), so the neighbouring tiles
// that share a line must run on one XCD: the linear id is remapped so that XCD x works
// through one contiguous eighth of the tiles, in order.  A bijection of the grid.
constexpr unsigned STZ_NXCD = 8;
__device__ __forceinline__ uint3 xcd_block()
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned T = gx * gy * gridDim.z;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned x = L % STZ_NXCD, i = L / STZ_NXCD, q = T / STZ_NXCD, r = T % STZ_NXCD;
    const unsigned Lp = x * q + (x < r ? x : r) + i;
    return make_uint3(Lp % gx, (Lp / gx) % gy, Lp / (gx * gy));
}

// ---------------------------------------------------------------------------------
// k_ypass: complex FFT along y (stride hp) for a tile of W k0-columns, in place
// ---------------------------------------------------------------------------------
template <typename R, int L1, int W, int PL = 0>
__global__ __launch_bounds__(NT_SOLVE) void k_ypass(cpx<R> *__restrict__ spec_a, const cpx<R> *__restrict__ twg,
                                                    SolveConsts k, R sgn, const SolveRep *reps)
{
    const uint3 bk = xcd_block();
    const View<R> vw = view_of<R>(reps, bk.z / k.C, k, nullptr, spec_a, nullptr, nullptr, nullptr);
    if (gated_off_p(vw.gate)) return;
    cpx<R> *__restrict__ spec = vw.spec;
    __shared__ cpx<R> buf[L1 * W];
    __shared__ cpx<R> tw[L1];
    const int tid = threadIdx.x;
    for (int i = tid; i < L1; i += NT_SOLVE) tw[i] = twg[i];
    const int k0base = bk.x * W;
    const int i2 = bk.y;
    const int c = bk.z % k.C;
    cpx<R> *base = spec + ((long long)c * k.rows + (long long)i2 * L1) * k.hp + k0base;
    // line = column of the tile; the first pass reads the columns from the spectrum and
    // the last pass writes them back
    fft_lines_io<R, L1, W, W, NT_SOLVE, true, true, PL>(
        buf, tw, sgn,
        [&](int col, int p) {
            return k0base + col < k.h0 ? gload(base + (long long)p * k.hp + col) : cpx<R>{0, 0};
        },
        [&](int col, int p, const cpx<R> &v) {
            if (k0base + col < k.h0) gstore(base + (long long)p * k.hp + col, v);
        });
}

// Gamma-hat on every wavevector of a fused tile (LZ points x C*W lines in LDS, line =
// c*W + col; 3-D: at y index i1): IsotropicMul_C, HadamardProdnSum_C x3,
// HadamardProdnSSum_C x2, HadamardProd2nSSum_C (STZ/elastic.c:802-809), with the
// average over the mirror bin's unit vector on Nyquist bins
template <typename R, int LZ, int W, int C, int DIM, int NTF>
__device__ __forceinline__ void gamma_tile(cpx<R> *buf, int k0base, int i1, const SolveConsts &k)
{
    constexpr int NL = C * W;
    const R lam = (R)k.lambda, mu = (R)k.mu, s0 = (R)k.s0, sh = (R)k.sh, sa = (R)k.sa;
    for (int kp = threadIdx.x; kp < LZ * W; kp += NTF) {
        const int col = kp % W, p = kp / W;
        int idx[3];
        idx[0] = k0base + col;
        if constexpr (DIM == 3) { idx[1] = i1; idx[2] = p; }
        else { idx[1] = p; idx[2] = 0; }
        double gd[3];
        unit_k<DIM>(idx, k.rn, gd);
        bool nyq = false;
#pragma unroll
        for (int a = 0; a < DIM; ++a) nyq |= (2 * idx[a] == k.n[a]);
        R g[3] = {(R)gd[0], (R)gd[1], (R)gd[2]};
        R gg[6];
        products<R, DIM>(g, gg);
        // the real parts, then the imaginary parts (Gamma-hat acts on each alone): half
        // the live values of both at once, which set the kernel's register count
        R *bufr = reinterpret_cast<R *>(buf);
#pragma unroll 1
        for (int part = 0; part < 2; ++part) {
            R e[C], o[C];
#pragma unroll
            for (int c = 0; c < C; ++c) e[c] = bufr[2 * (p * NL + c * W + col) + part];
            gamma_part<R, C>(e, g, gg, lam, mu, s0, sh, sa, o);
            if (nyq) {
                // average with the mirror bin's unit vector (Nyquist components negated)
                R gm[3] = {g[0], g[1], g[2]};
#pragma unroll
                for (int a = 0; a < DIM; ++a)
                    if (2 * idx[a] == k.n[a]) gm[a] = -gm[a];
                R ggm[6];
                products<R, DIM>(gm, ggm);
                R o2[C];
                gamma_part<R, C>(e, gm, ggm, lam, mu, s0, sh, sa, o2);
#pragma unroll
                for (int c = 0; c < C; ++c) o[c] = (o[c] + o2[c]) * (R)0.5;
            }
#pragma unroll
            for (int c = 0; c < C; ++c) bufr[2 * (p * NL + c * W + col) + part] = o[c];
        }
    }
}

// ---------------------------------------------------------------------------------
// k_fused: last axis forward FFT, Gamma-hat, inverse FFT for one tile of W k0-columns
// (3-D: at a fixed y index; 2-D: the y axis itself is the last axis).  The
// tile of all C components (LZ x C*W complex values) stays in LDS from load to store:
// the C*W lines are transformed together, Gamma-hat mixes the C components of each of
// the LZ*W wavevectors in place, and the inverse transform follows.
// ---------------------------------------------------------------------------------
template <typename R, int LZ, int W, int C, int DIM, int NTF, int OP = 0, int PL = 0>
__global__ __launch_bounds__(NTF) void k_fused(cpx<R> *__restrict__ spec_a, const cpx<R> *__restrict__ twg,
                                               SolveConsts k, const SolveRep *reps)
{
    constexpr int NL = C * W;             // lines, component-major: line = c*W + col
    constexpr int NE = LZ * NL;
    const uint3 bk = xcd_block();
    const View<R> vw = view_of<R>(reps, bk.z, k, nullptr, spec_a, nullptr, nullptr, nullptr);
    if (gated_off_p(vw.gate)) return;
    cpx<R> *__restrict__ spec = vw.spec;
    __shared__ cpx<R> buf[NE];
    __shared__ cpx<R> tw[LZ];
    const int tid = threadIdx.x;
    for (int i = tid; i < LZ; i += NTF) tw[i] = twg[i];
    const int k0base = bk.x * W;
    const int i1 = (DIM == 3) ? bk.y : 0;
    // element (p, col) of component c lives at spec[(c*rows + p*pstride + i1) * hp + k0]
    const long long pstride = (DIM == 3) ? k.n[1] : 1;
    auto gaddr = [&](int line, int p) {
        return spec + ((long long)(line / W) * k.rows + (long long)p * pstride + i1) * k.hp + k0base + line % W;
    };
    // forward: the first pass reads the tile from the spectrum, the result stays in LDS
    fft_lines_io<R, LZ, NL, NL, NTF, true, false, PL>(
        buf, tw, (R)-1,
        [&](int line, int p) { return k0base + line % W < k.h0 ? gload(gaddr(line, p)) : cpx<R>{0, 0}; },
        [](int, int, const cpx<R> &) {});
    // Gamma-hat (OP 0), the displacement operator (OP 1) or the heat multiplier (OP 2) on
    // each wavevector of the tile
    const R lam = (R)k.lambda, mu = (R)k.mu, s0 = (R)k.s0, sh = (R)k.sh, sa = (R)k.sa;
    if constexpr (OP == 1) {
        // u_hat_a = (tmp_a(im)*(1/mu)/|k|, -1.0*tmp_a(re)*(1/mu)/|k|) (HadamardSiDivd,
        // STZ/elastic.c:754-763) into component slots 0..DIM-1; on Nyquist bins the
        // average with the mirror unit vector, as Re() of the full transform applies
        const R nalpha = (R)(-1.0 * k.alpha);
        for (int kp = tid; kp < LZ * W; kp += NTF) {
            const int col = kp % W, p = kp / W;
            int idx[3];
            idx[0] = k0base + col;
            if constexpr (DIM == 3) { idx[1] = i1; idx[2] = p; }
            else { idx[1] = p; idx[2] = 0; }
            double gd[3];
            const R kmag = (R)unit_k<DIM>(idx, k.rn, gd);
            bool nyq = false;
#pragma unroll
            for (int a = 0; a < DIM; ++a) nyq |= (2 * idx[a] == k.n[a]);
            const R g[3] = {(R)gd[0], (R)gd[1], (R)gd[2]};
            R er[C], ei[C], pr[3], pi[3];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const cpx<R> x = buf[p * NL + c * W + col];
                er[c] = x.x;
                ei[c] = x.y;
            }
            disp_part<R, C, DIM>(er, g, lam, mu, nalpha, pr);
            disp_part<R, C, DIM>(ei, g, lam, mu, nalpha, pi);
            if (nyq) {
                R gm[3] = {g[0], g[1], g[2]};
#pragma unroll
                for (int a = 0; a < DIM; ++a)
                    if (2 * idx[a] == k.n[a]) gm[a] = -gm[a];
                R pr2[3], pi2[3];
                disp_part<R, C, DIM>(er, gm, lam, mu, nalpha, pr2);
                disp_part<R, C, DIM>(ei, gm, lam, mu, nalpha, pi2);
#pragma unroll
                for (int a = 0; a < DIM; ++a) {
                    pr[a] = (pr[a] + pr2[a]) * (R)0.5;
                    pi[a] = (pi[a] + pi2[a]) * (R)0.5;
                }
            }
#pragma unroll
            for (int a = 0; a < DIM; ++a) buf[p * NL + a * W + col] = {pi[a] * s0 / kmag, (R)-1.0 * pr[a] * s0 / kmag};
        }
    } else if constexpr (OP == 2) {
        // HeatConducting's multiplier exp(-g2*dt), g2 = |k| of the reference grid and 1 at
        // k = 0 (STZ/elastic.c:563-566); even in k, so the half spectrum needs no averaging
        for (int kp = tid; kp < LZ * W; kp += NTF) {
            const int col = kp % W, p = kp / W;
            int idx[3];
            idx[0] = k0base + col;
            if constexpr (DIM == 3) { idx[1] = i1; idx[2] = p; }
            else { idx[1] = p; idx[2] = 0; }
            double gd[3];
            const double g2 = unit_k<DIM>(idx, k.rn, gd);
            const R m = (R)exp(-g2 * k.hdt);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const cpx<R> x = buf[p * NL + c * W + col];
                buf[p * NL + c * W + col] = {x.x * m, x.y * m};
            }
        }
    } else {
        gamma_tile<R, LZ, W, C, DIM, NTF>(buf, k0base, i1, k);
    }
    __syncthreads();
    // inverse from the tile; the last pass writes the spectrum
    fft_lines_io<R, LZ, NL, NL, NTF, false, true, PL>(
        buf, tw, (R)1, [](int, int) { return cpx<R>{0, 0}; },
        [&](int line, int p, const cpx<R> &v) {
            if (k0base + line % W < k.h0) gstore(gaddr(line, p), v);
        });
}

// k0 columns per tile for the strided passes
constexpr int cols_for(int L) { return L >= 512 ? 4 : (L >= 256 ? 8 : 16); }
// k0 columns of the fused tile: the largest power of two (2..16) keeping the
// LZ x C*W tile within 96 KB of LDS
template <typename R, int C>
constexpr int fused_cols_for(int L)
{
    int w = 16;
    while (w > 2 && (long long)L * C * w * (long long)sizeof(cpx<R>) > 98304) w /= 2;
    return w;
}
constexpr int fused_threads(int elems) { return elems >= 4096 ? 1024 : (elems >= 2048 ? 512 : 256); }

// Tile variants the autotuner may pick (compiled for every length): row counts
// NR in [2, 2*rows_for] for the x passes with a tile of at most 64 KB (the default may
// be below rows_for when the row total is not a multiple of it), y-pass
// widths W in [cols_for/2, 2*cols_for] with at most 64 KB, fused widths W in
// [fused_cols_for/4, fused_cols_for].
template <typename R>
constexpr bool y_variant(int L, int W)
{
    return W >= 4 && W * 2 >= cols_for(L) && W <= 2 * cols_for(L) && (long long)L * W * (long long)sizeof(cpx<R>) <= 65536;
}
template <typename R, int C>
constexpr bool f_variant(int L, int W)
{
    return W >= 2 && W * 4 >= fused_cols_for<R, C>(L) && W <= fused_cols_for<R, C>(L);
}


// a tuning value is a tile (rows or columns) in its low byte and a radix plan above it
constexpr int tune_tile(int v) { return v & 255; }
constexpr int tune_plan(int v) { return (v >> 8) & 15; }

// launch constants of a context (the batched path overrides the per-replica fields)
inline SolveConsts solve_consts(const stz_ctx *ctx)
{
    SolveConsts k{};
    k.gate = nullptr;
    k.lambda = ctx->k.lambda;
    k.mu = ctx->k.mu;
    k.alpha = ctx->k.alpha;
    k.s0 = 1.0 / ctx->k.mu;
    k.sh = 0.5 / ctx->k.mu;
    k.sa = -1.0 * ctx->k.alpha / ctx->k.mu;
    k.invN = 1.0 / (double)ctx->N;
    for (int c = 0; c < 6; ++c) k.epsavg[c] = ctx->epsavg[c];
    for (int a = 0; a < 3; ++a) {
        k.n[a] = ctx->n[a];
        k.rn[a] = 1.0 / (double)ctx->n[a];
    }
    k.N = ctx->N;
    k.rows = ctx->n[1] * ctx->n[2];
    k.h0 = ctx->h0;
    k.hp = ctx->hp;
    k.C = ctx->C;
    k.L[0] = ctx->p.lx;
    k.L[1] = ctx->p.ly;
    k.L[2] = ctx->p.lz;
    return k;
}

}  // namespace

// the launch sequence of each storage precision (stz_solve_f64.hip, stz_solve_f32.hip);
// the first argument only selects the precision
int solve_entry(double, stz_ctx *ctx, int gated, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts);
int solve_entry(float, stz_ctx *ctx, int gated, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts);
int spectral_entry(double, stz_ctx *ctx);
int spectral_entry(float, stz_ctx *ctx);
int disp_entry(double, stz_ctx *ctx, double *u);
int disp_entry(float, stz_ctx *ctx, double *u);
