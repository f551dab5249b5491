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
    long long N;
};

// one source of an update: its grid coordinates and stored eigenstrain increment
template <typename R>
struct LocalSrc {
    int c[3];
    R d[STZ_MAXS];
};

template <typename R, int DIM>
__device__ __forceinline__ void load_src(const LocalGeom &g, const R *deps_last, long long v, LocalSrc<R> &s)
{
    constexpr int S = DIM == 2 ? 2 : 5;
    s.c[0] = (int)(v % g.n[0]);
    s.c[1] = (int)((v / g.n[0]) % g.n[1]);
    s.c[2] = (int)(v / ((long long)g.n[0] * g.n[1]));
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
        const long long jj = ((long long)d2 * g.n[1] + d1) * g.n[0] + d0;
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

// Updates the voxels [blk*span, min(N, (blk+1)*span)) of one workgroup: sigma += uni +
// sum over sources, and writes the workgroup's FP64 partial sums of the new stress to
// spart[blk][c].  Sources must be in `src` (shared memory) already.
template <typename R, int DIM>
__device__ void local_block(const LocalGeom &g, const R *__restrict__ K, const LocalSrc<R> *src, int ns,
                           const double *uni, R *sigma, long long blk, long long span, double *spart, int *nanflag,
                           double *red)
{
    constexpr int C = DIM == 2 ? 3 : 6;
    const int tid = threadIdx.x;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    int bad = 0;
    const long long v0 = blk * span;
    const long long v1 = v0 + span < g.N ? v0 + span : g.N;
    for (long long v = v0 + tid; v < v1; v += NT_LOCAL) {
        const int i0 = (int)(v % g.n[0]);
        const int i1 = (int)((v / g.n[0]) % g.n[1]);
        const int i2 = (int)(v / ((long long)g.n[0] * g.n[1]));
        R ds[C];
        local_sum<R, DIM>(g, K, src, ns, i0, i1, i2, ds);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const double s = ((double)sigma[(long long)c * g.N + v] + uni[c]) + (double)ds[c];
            bad |= (s != s);
            sigma[(long long)c * g.N + v] = (R)s;
            acc[c] += s;
        }
    }
    if (bad) atomicOr(nanflag, 1);
    for (int c = 0; c < C; ++c) {
        red[tid] = acc[c];
        __syncthreads();
        for (int w = NT_LOCAL / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] = red[tid] + red[tid + w];
            __syncthreads();
        }
        if (tid == 0) spart[blk * C + c] = red[0];
        __syncthreads();
    }
}

}  // namespace
