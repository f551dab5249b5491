// stz_local_dev.h -- event-local stress update, the device side.
//
// The stress solve (STZ/elastic.c:779-828 behind the uniform-moduli wrapper
// STZ/elastic.c:977-1066) is linear in the eigenstrain field eps0 and in the mean strain
// epsavg, and its Green's operator is a Fourier multiplier, i.e. a periodic convolution.
// A transform at voxel s adds a traceless tensor d (S stored components) to eps0 there
// only, so the re-solved stress is
//
//     sigma_new(x) = sigma_solved_before(x) + C0:(epsavg_now - epsavg_before)
//                    + sum_k d_k K_k((x - s) mod n),
//
// with K_k the stress response to the k-th basis eigenstrain placed at voxel 0 (all C
// components, including the local -C0:eps0 term), built once per grid and elastic
// constants by FP64 FFT solves of the S unit impulses.
//
// The reference keeps Re() of a full complex inverse transform (STZ/elastic.c:811-824),
// which applies the average of the multiplier at k and -k: the operator is even, and
// K(-d) = K(d) in every component.  So K is stored for the half space of offsets whose
// slowest coordinate is in [0, n/2] (the first n0*n1*(n2/2+1) voxels of the grid), and an
// offset outside it reads K(-d).  (Single-axis reflections are not symmetries: the
// reference's round() k vectors, STZ/init.c:126-129, put every Nyquist component at -pi.)
// A workgroup takes one row of offsets (d1, d2) of the half space and updates the two
// grid rows s + (., d1, d2) and s - (., d1, d2), which need the same response row, so each
// response value is read once per update.  A 2-D grid is treated as n0 x 1 x n1.
//
// sigma_solved_before is the field minus the uniform loading increments made since the
// last solve: in FP64 they are in the field (usum), in mixed precision they are kept as
// offsets outside it.
#pragma once

#include "stz_internal.h"


namespace {

// voxels along x per half-space row group of one workgroup (each updates two grid rows),
// FP64 and FP32 storage (measured at 128^3: 256/512/1024 within 2 % in FP64, 1024 7 %
// faster than 512 in FP32)
#ifndef STZ_LOCAL_ROW_VOXELS
#define STZ_LOCAL_ROW_VOXELS 512
#endif
#ifndef STZ_LOCAL_ROW_VOXELS_FP32
#define STZ_LOCAL_ROW_VOXELS_FP32 1024
#endif
__host__ __device__ constexpr int local_row_voxels(bool fp32)
{
    return fp32 ? STZ_LOCAL_ROW_VOXELS_FP32 : STZ_LOCAL_ROW_VOXELS;
}

struct LocalGeom {
    int n[3];          // 3-D: the grid; 2-D: (n0, 1, n1)
    int N;             // below 2^31 (checked when the response is built)
    int h2;            // n[2]/2 + 1: extent of the stored half space along the slowest axis
    int H;             // n[0]*n[1]*h2 values per response field
    int ipw;           // offset rows per workgroup (several when a row is short)
    int items;         // offset rows n[1]*h2
};

__host__ __device__ inline LocalGeom local_geom(int dim, const int *n, int row_voxels)
{
    LocalGeom g;
    g.n[0] = n[0];
    g.n[1] = dim == 2 ? 1 : n[1];
    g.n[2] = dim == 2 ? n[1] : n[2];
    g.N = g.n[0] * g.n[1] * g.n[2];
    g.h2 = g.n[2] / 2 + 1;
    g.H = g.n[0] * g.n[1] * g.h2;
    g.ipw = g.n[0] >= row_voxels ? 1 : row_voxels / g.n[0];
    g.items = g.n[1] * g.h2;
    return g;
}

// one source of an update: its grid coordinates (kernel geometry) and stored
// eigenstrain increment
template <typename R>
struct LocalSrc {
    int c[3];
    R d[STZ_MAXS];
};

template <typename R, int DIM>
__device__ __forceinline__ void load_src(const LocalGeom &g, const R *deps_last, int v, LocalSrc<R> &s)
{
    constexpr int S = DIM == 2 ? 2 : 5;
    s.c[0] = v % g.n[0];
    s.c[1] = (v / g.n[0]) % g.n[1];
    s.c[2] = v / (g.n[0] * g.n[1]);
#pragma unroll
    for (int k = 0; k < S; ++k) s.d[k] = deps_last[(long long)k * g.N + v];
}

// The response table, compacted by the elastic operator's major symmetry.  K_kc(d), the
// stress component c at offset d of a unit stored strain component k (a traceless basis
// tensor t_k), is G(d) : t_k with G the operator from eigenstrain to stress, which has
// G_ijkl = G_klij.  So the S x S matrix A_ik = t_i : G : t_k (the response's work-like
// contraction, shear entries weighted 2) is symmetric, and K is fixed by A and by one
// normal stress row: the table stores, per offset, the S(S+1)/2 entries of A (upper
// triangle, row by row) and the S responses of the reference normal component Z (zz in
// 3-D, yy in 2-D): 20 values instead of 30 in 3-D, 5 instead of 6 in 2-D.  With u = A d and
// z = sum_k d_k K_kZ for a strain increment d: sigma_Z = z, the other normal components
// u_i + z (A's normal rows are t_i = e_i - e_Z), the shear components u_i / 2.
template <int DIM>
struct Resp {
    static constexpr int S = DIM == 2 ? 2 : 5, C = DIM == 2 ? 3 : 6;
    static constexpr int NSYM = S * (S + 1) / 2, P = NSYM + S;   // planes of the table
    static constexpr int Z = DIM == 2 ? 1 : 2;                   // reference normal component
    static constexpr int NN = DIM == 2 ? 1 : 2;                  // normal rows of A
    // the stress component of row i of A
    __host__ __device__ static constexpr int comp(int i) { return DIM == 2 ? (i == 0 ? 0 : 2) : (i < 2 ? i : i + 1); }
    // plane of A_ik
    __host__ __device__ static constexpr int sym(int i, int k)
    {
        return i <= k ? i * S - i * (i - 1) / 2 + (k - i) : k * S - k * (k - 1) / 2 + (i - k);
    }
};

// da[c] = sum_k d_k K_kc from the P table values kv of one offset
template <typename R, int DIM>
__device__ __forceinline__ void resp_apply(const R *kv, const R *d, R *da)
{
    using Q = Resp<DIM>;
    R u[Q::S];
#pragma unroll
    for (int i = 0; i < Q::S; ++i) {
        u[i] = (R)0;
#pragma unroll
        for (int k = 0; k < Q::S; ++k) u[i] = u[i] + d[k] * kv[Q::sym(i, k)];
    }
    R z = (R)0;
#pragma unroll
    for (int k = 0; k < Q::S; ++k) z = z + d[k] * kv[Q::NSYM + k];
    da[Q::Z] = z;
#pragma unroll
    for (int i = 0; i < Q::S; ++i) da[Q::comp(i)] = i < Q::NN ? u[i] + z : u[i] * (R)0.5;
}

// index of offset d (each coordinate in [0, n)) in the half-space table: d itself when
// its slowest coordinate is at most n2/2, else -d
__device__ __forceinline__ int half_index(const LocalGeom &g, int d0, int d1, int d2)
{
    if (2 * d2 > g.n[2]) {
        d0 = d0 ? g.n[0] - d0 : 0;
        d1 = d1 ? g.n[1] - d1 : 0;
        d2 = g.n[2] - d2;
    }
    return (d2 * g.n[1] + d1) * g.n[0] + d0;
}

// C0:e for the uniform strain change (the form of the solve's epilogue)
template <int C>
__host__ __device__ inline void c0_mul(const double *e, double lam, double mu, double *s)
{
    if constexpr (C == 3) {
        const double tr = e[0] + e[1];
        s[0] = lam * tr + 2 * mu * e[0];
        s[1] = lam * tr + 2 * mu * e[1];
        s[2] = 2 * mu * e[2];
    } else {
        const double tr = (e[0] + e[1]) + e[2];
        for (int c = 0; c < 3; ++c) s[c] = lam * tr + 2 * mu * e[c];
        for (int c = 3; c < 6; ++c) s[c] = 2 * mu * e[c];
    }
}

constexpr int NT_LOCAL = 256;
constexpr int NW_LOCAL = NT_LOCAL / 64;

__device__ __forceinline__ double wsum_local(double x)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) x = x + __shfl_xor(x, d, 64);
    return x;
}

// the C stress components of voxel v
template <typename R, int C>
__device__ __forceinline__ void local_load(const LocalGeom &g, const R *sigma, int v, R *sv)
{
#pragma unroll
    for (int c = 0; c < C; ++c) sv[c] = gld_stream<4>(sigma + (long long)c * g.N + v);
}

// sigma[v] = sv + uni + ds (FP64 sum, rounded to the storage precision once); sv holds
// the components local_load read before any store of the workgroup's rows
template <typename R, int C>
__device__ __forceinline__ void local_store(const LocalGeom &g, R *sigma, int v, const R *sv, const double *uni,
                                            const R *ds, double *acc, int &bad)
{
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double s = ((double)sv[c] + uni[c]) + (double)ds[c];
        bad |= (s != s);
        gstore(sigma + (long long)c * g.N + v, (R)s);
        acc[c] += s;
    }
}

// sum over sources of sum_k d_k K_k(x - s) at voxel (x0, x1, x2) (any number of sources)
template <typename R, int DIM>
__device__ __forceinline__ void local_sum(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src, int ns,
                                          int x0, int x1, int x2, R *out)
{
    constexpr int C = DIM == 2 ? 3 : 6;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = (R)0;
    for (int j = 0; j < ns; ++j) {
        int d0 = x0 - src[j].c[0], d1 = x1 - src[j].c[1], d2 = x2 - src[j].c[2];
        if (d0 < 0) d0 += g.n[0];
        if (d1 < 0) d1 += g.n[1];
        if (d2 < 0) d2 += g.n[2];
        const int hj = half_index(g, d0, d1, d2);
        R kv[Resp<DIM>::P], da[C];   // the table values of the offset, loaded before their use
#pragma unroll
        for (int i = 0; i < Resp<DIM>::P; ++i) kv[i] = gload(K + (long long)i * g.H + hj);
        resp_apply<R, DIM>(kv, src[j].d, da);
#pragma unroll
        for (int c = 0; c < C; ++c) out[c] = out[c] + da[c];
    }
}

// One workgroup: the offset rows w*ipw .. of the half space (relative to source 0) and,
// for each, the grid rows s0 + (., d1, d2) and (off the boundary planes d2 = 0, n2/2)
// s0 - (., d1, d2).  With one source a thread reads the response at (d0, d1, d2) once
// and applies it to x = s0 + d and x = s0 - d.  Writes the workgroup's FP64 partial sums
// of the new stress, in a fixed order, to spart[w][c]; wred holds NW_LOCAL*6 doubles.
template <typename R, int DIM>
__device__ void local_rows(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src, int ns,
                           const double *uni, R *sigma, int w, double *spart, int *nanflag, double *wred)
{
    constexpr int C = DIM == 2 ? 3 : 6;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    int bad = 0;
    // rows are laid out around source 0 (any origin partitions the grid; with no source,
    // voxel 0 and only the uniform part)
    LocalSrc<R> s0;
    if (ns > 0) {
        s0 = src[0];
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) s0.c[a] = 0;
#pragma unroll
        for (int k2 = 0; k2 < STZ_MAXS; ++k2) s0.d[k2] = (R)0;
    }
    for (int e = tid; e < g.ipw * n0; e += NT_LOCAL) {
        const int item = w * g.ipw + e / n0;
        if (item >= g.items) break;
        const int d0 = e % n0;
        const int d1 = item % n1, d2 = item / n1;
        const bool pair = d2 != 0 && 2 * d2 != n2;
        const int a1 = (s0.c[1] + d1) % n1, a2 = (s0.c[2] + d2) % n2;
        const int b1 = (s0.c[1] - d1 + n1) % n1, b2 = (s0.c[2] - d2 + n2) % n2;
        const int va = (a2 * n1 + a1) * n0 + (s0.c[0] + d0) % n0;
        const int vb = (b2 * n1 + b1) * n0 + (s0.c[0] - d0 + n0) % n0;
        // va != vb on a pair row, and every voxel belongs to one row of one workgroup:
        // the stress of both voxels is read before anything is written, so the loads are
        // all in flight together instead of one load-wait-store round trip per component
        R sa[C], sb[C];
        R da[C], db[C];
        if (ns == 1) {
            // the response row first, then the stress: all S*C + 2*C loads are issued
            // before the first use
            const int hj = (d2 * n1 + d1) * n0 + d0;
            R kv[Resp<DIM>::P];
#pragma unroll
            for (int i = 0; i < Resp<DIM>::P; ++i) kv[i] = gld_stream<4>(K + (long long)i * g.H + hj);
            local_load<R, C>(g, sigma, va, sa);
            if (pair) local_load<R, C>(g, sigma, vb, sb);
            resp_apply<R, DIM>(kv, s0.d, da);
#pragma unroll
            for (int c = 0; c < C; ++c) db[c] = da[c];
        } else {
            local_load<R, C>(g, sigma, va, sa);
            if (pair) local_load<R, C>(g, sigma, vb, sb);
            local_sum<R, DIM>(g, K, src, ns, va % n0, a1, a2, da);
            if (pair) local_sum<R, DIM>(g, K, src, ns, vb % n0, b1, b2, db);
        }
        local_store<R, C>(g, sigma, va, sa, uni, da, acc, bad);
        if (pair) local_store<R, C>(g, sigma, vb, sb, uni, db, acc, bad);
    }
    if (bad) atomicOr(nanflag, 1);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double x = wsum_local(acc[c]);
        if (lane == 0) wred[wv * 6 + c] = x;
    }
    __syncthreads();
    if (tid < C) {
        double t = wred[tid];
        for (int i = 1; i < NW_LOCAL; ++i) t = t + wred[i * 6 + tid];
        spart[(long long)w * C + tid] = t;
    }
}

}  // namespace
