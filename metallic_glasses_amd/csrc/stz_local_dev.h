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
// constants by FP64 FFT solves of the S unit impulses and kept for the whole grid.
// (The reference's k vectors use round(), STZ/init.c:126-129, which puts every Nyquist
// component at -pi; on wavevectors with two or more Nyquist components the operator is
// then not symmetric under the reflection of one axis, so K cannot be folded to one
// octant of offsets without changing the result.)
//
// sigma_solved_before is the field minus the uniform loading increments made since the
// last solve: in FP64 they are in the field (usum), in mixed precision they are kept as
// offsets outside it.
#pragma once

#include "stz_internal.h"


namespace {

struct LocalGeom {
    int n[3];
    int N;             // below 2^31 (checked when the response is built)
};

// one source of an update: its grid coordinates and stored eigenstrain increment
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

// sum over sources of sum_k d_k K_k((x - s) mod n) for the C components at voxel (i0, i1, i2)
template <typename R, int DIM>
__device__ __forceinline__ void local_sum(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src, int ns,
                                          int i0, int i1, int i2, R *out)
{
    constexpr int C = DIM == 2 ? 3 : 6, S = DIM == 2 ? 2 : 5;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = (R)0;
    for (int s = 0; s < ns; ++s) {
        int d0 = i0 - src[s].c[0], d1 = i1 - src[s].c[1], d2 = i2 - src[s].c[2];
        if (d0 < 0) d0 += g.n[0];
        if (d1 < 0) d1 += g.n[1];
        if (d2 < 0) d2 += g.n[2];
        const int jj = (d2 * g.n[1] + d1) * g.n[0] + d0;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const R dk = src[s].d[k];
#pragma unroll
            for (int c = 0; c < C; ++c) out[c] = out[c] + dk * K[((long long)k * C + c) * g.N + jj];
        }
    }
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
constexpr int VPT_LOCAL = 4;                          // voxels per thread
constexpr int SPAN_LOCAL = NT_LOCAL * VPT_LOCAL;      // voxels per workgroup
constexpr int NW_LOCAL = NT_LOCAL / 64;

__device__ __forceinline__ double wsum_local(double x)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) x = x + __shfl_xor(x, d, 64);
    return x;
}

// one voxel: sigma += uni + the shifted responses of the sources; returns the new value
// of each component in FP64 (before rounding to the storage precision)
template <typename R, int DIM, bool ONE>
__device__ __forceinline__ void local_voxel(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src,
                                            const LocalSrc<R> &s0, int ns, const double *uni, R *sigma, int v,
                                            double *acc, int &bad)
{
    constexpr int C = DIM == 2 ? 3 : 6;
    const int i0 = v % g.n[0];
    const int r = v / g.n[0];
    const int i1 = r % g.n[1];
    const int i2 = r / g.n[1];
    R ds[C];
    if constexpr (ONE) local_sum<R, DIM>(g, K, &s0, 1, i0, i1, i2, ds);
    else local_sum<R, DIM>(g, K, src, ns, i0, i1, i2, ds);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double s = ((double)sigma[(long long)c * g.N + v] + uni[c]) + (double)ds[c];
        bad |= (s != s);
        sigma[(long long)c * g.N + v] = (R)s;
        acc[c] += s;
    }
}

// Updates the voxels [blk*SPAN_LOCAL, ...) of one workgroup (VPT_LOCAL per thread,
// consecutive lanes on consecutive voxels) and writes the workgroup's FP64 partial sums
// of the new stress, in a fixed order, to spart[blk][c].  Sources must be in `src`
// (shared memory) already; wred holds NW_LOCAL*6 doubles.
template <typename R, int DIM>
__device__ void local_block(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src, int ns,
                           const double *uni, R *sigma, long long blk, double *spart, int *nanflag, double *wred)
{
    constexpr int C = DIM == 2 ? 3 : 6;
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    int bad = 0;
    const int v0 = (int)blk * SPAN_LOCAL;
    if (ns == 1) {
        const LocalSrc<R> s0 = src[0];
#pragma unroll
        for (int t = 0; t < VPT_LOCAL; ++t) {
            const int v = v0 + t * NT_LOCAL + tid;
            if (v < g.N) local_voxel<R, DIM, true>(g, K, src, s0, 1, uni, sigma, v, acc, bad);
        }
    } else {
        for (int t = 0; t < VPT_LOCAL; ++t) {
            const int v = v0 + t * NT_LOCAL + tid;
            if (v < g.N) local_voxel<R, DIM, false>(g, K, src, src[0], ns, uni, sigma, v, acc, bad);
        }
    }
    if (bad) atomicOr(nanflag, 1);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double x = wsum_local(acc[c]);
        if (lane == 0) wred[w * 6 + c] = x;
    }
    __syncthreads();
    if (tid < C) {
        double t = wred[tid];
        for (int i = 1; i < NW_LOCAL; ++i) t = t + wred[i * 6 + tid];
        spart[blk * C + tid] = t;
    }
}

}  // namespace
