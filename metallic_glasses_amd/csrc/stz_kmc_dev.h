// stz_kmc_dev.h -- device code shared by the per-call kernels (stz_rates.hip) and the
// resident batched step (stz_resident.hip): the per-voxel physics of SofteningFactor and
// ActEnergy (STZ/init.c:561-739) and of the KMC rates (STZ/evolution.c:120-128), the
// deterministic wavefront/block scans of the residence-time selection
// (STZ/evolution.c:136-224), the athermal compaction (STZ/evolution.c:52-77) and the glass
// branch of VoxelShearTransform (STZ/elastic.c:361-520).  Every caller evaluates the same
// inline code, so the per-call and the resident paths produce identical bits.
#pragma once

#include "stz_internal.h"

namespace {

// modes whose operands the rate pass loads ahead of the one it evaluates: measured per
// storage precision at 512x512x256 and 256^3 (FP32 2: -5 % / -3 % against 1; FP64 3: -5 % /
// -2 %); STZ_RATE_DEPTH overrides both
template <typename R>
constexpr int rate_depth()
{
#ifdef STZ_RATE_DEPTH
    return STZ_RATE_DEPTH;
#else
    return sizeof(R) == 4 ? 2 : 3;
#endif
}

constexpr int NT = 256;
static_assert(NT == STZ_CHUNK, "a chunk is one workgroup's scan");
constexpr int WAVE = 64;
constexpr int NW = NT / WAVE;

template <typename R>
struct RateArgs {
    R *sigma;
    const R *q0, *deps, *deps_last;
    R *q0eff;                  // cached Q0*exp(-D*S) [M][N]
    R *scache;                 // S the cache entry was computed with [N] (NaN: invalid)
    const double *perm, *temp, *tlast;
    R *minq;
    uint8_t *argm;
    const uint8_t *mro;        // crystal-like MRO types (nullptr: glass everywhere)
    double *bsum;
    int *acnt;                 // per voxel block: voxels meeting the athermal criterion
    unsigned long long *nneg;  // Q<0 counter (exact integer atomics)
    double thr;
    long long N;
    long long nblk;
    int M;
    int softening;
    int pend_comp;
    double pend_delta;
    double soff[6];
    double totime, tTemp, kT, ratepre, voxel, lnpre;
};

// ---------------------------------------------------------------------------------
// per-voxel physics, evaluated with every operation rounded separately so that any
// two call sites produce identical bits
// ---------------------------------------------------------------------------------
#pragma clang fp contract(off)

template <int DIM>
struct Comp {
    static constexpr int C = DIM == 2 ? 3 : 6;
    static constexpr int S = DIM == 2 ? 2 : 5;
};

// stored S components -> C tensor components (yy = -xx in 2-D; zz = -(xx + yy) in 3-D)
template <typename R, int DIM>
__device__ __forceinline__ void expand(const R *st, R *t)
{
    if constexpr (DIM == 2) {
        t[0] = st[0]; t[1] = -st[0]; t[2] = st[1];
    } else {
        t[0] = st[0]; t[1] = st[1]; t[2] = -(st[0] + st[1]); t[3] = st[2]; t[4] = st[3]; t[5] = st[4];
    }
}

// a:b with weights (1,1,2) or (1,1,1,2,2,2), left to right (STZ/init.c:585-586)
template <typename R, int DIM>
__device__ __forceinline__ R contract(const R *a, const R *b)
{
    constexpr int C = Comp<DIM>::C;
    R acc = a[0] * b[0];
#pragma unroll
    for (int c = 1; c < C; ++c) {
        const bool two = (DIM == 2) ? (c == 2) : (c >= 3);
        acc = acc + (two ? ((R)2 * a[c]) * b[c] : a[c] * b[c]);
    }
    return acc;
}

template <typename R>
struct VoxelState {
    R sg[6];
    R pd[6];
    R norm2;
    R soft;
    const R *q0src; // the cached Q0*exp(-D*S) when fresh, else Q0 (selected once per voxel)
    R fac;          // crystal-like MRO: factor on the modifier and the softening per mode
    int t;          // MRO type (0 glass)
    bool fresh;     // the cached Q0*exp(-D*S) of this voxel is valid for this step's S
    bool settle;    // the entry is written as settled (its temporary softening term is 0)
};

// NO_SF_XTAL_MRO factors (STZ/init.c:645-667): six-fold 0.1, two-fold 0.01, ten-fold 0.9
template <typename R>
__device__ __forceinline__ R mro_factor(int t)
{
    return t == 1 ? (R)0.1 : (t == 2 ? (R)0.01 : (R)0.9);
}

// Softening S (STZ/init.c:731) and stress of one voxel.  Q0*exp(-D*S) (STZ/init.c:684)
// changes only when S changes or the voxel transforms, so it is cached per (mode,
// voxel) with the S it was computed for; S equals Perm bit for bit except shortly after
// a transform, so almost every voxel takes the cached value, which is the bit-identical
// result of the same expression.
//
// The entry scache[v] is "settled" when its sign bit is set (and it is not a NaN): then
// S = -scache[v], the temporary term Temp*exp((tlast - t)/tTemp) was exactly 0 when it
// was written and stays 0 (the clock only advances; a transform, an upload, a state
// initialisation, a checkpoint load or a clock reset invalidates the entry), and Perm,
// Temp and tlast are not read.  Otherwise S is computed and the entry is fresh when it
// holds that S.  An S with its sign bit set (a negative Perm upload) is never cached.
//
// LEAN: no cache at all (every voxel evaluates Q0*exp(-D*S) afresh, which gives the same
// bits as a cached entry); a.q0eff and a.scache are not touched.
template <typename R, int DIM, bool XTAL = false, bool LEAN = false>
__device__ __forceinline__ void load_voxel(const RateArgs<R> &a, long long v, bool apply_pending, VoxelState<R> &s)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    const R sc = LEAN ? (R)__builtin_nan("") : ld_stream<2>(a.scache + v);
    s.settle = false;
    if (!LEAN && __builtin_signbit(sc) && sc == sc) {
        s.soft = -sc;
        s.fresh = true;
    } else {
        double soft = 0.0;
        bool tzero = true;
        if (a.softening) {
            const double tt = ld_stream<2>(a.temp + v) * exp((ld_stream<2>(a.tlast + v) - a.totime) / a.tTemp);
            soft = ld_stream<2>(a.perm + v) + tt;                                            // :731
            tzero = (tt == 0.0);
        }
        s.soft = (R)soft;
        s.fresh = (sc == s.soft);                     // false for the NaN of an invalidated entry
        s.settle = tzero && !__builtin_signbit(s.soft);
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        R x = ld_stream<2>(a.sigma + (long long)c * a.N + v);
        if constexpr (sizeof(R) == 8) {
            if (apply_pending && c == a.pend_comp) {
                x = x + (R)a.pend_delta;      // Sigma[c][idx] += delta (STZ/evolution.c:311-314)
                a.sigma[(long long)c * a.N + v] = x;
            }
        } else {
            x = (R)((double)x + a.soff[c]);
        }
        s.sg[c] = x;
    }
    if constexpr (XTAL) {
        s.t = (int)a.mro[v];
        s.fac = s.t ? mro_factor<R>(s.t) : (R)1;
    } else {
        s.t = 0;
        s.fac = (R)1;
    }
    // one address select per voxel (a select between two struct fields inside the mode
    // loop is compiled through private memory)
    if constexpr (LEAN) s.fresh = false;
    const uintptr_t qa = (uintptr_t)a.q0eff, qb = (uintptr_t)a.q0;
    s.q0src = (const R *)(s.fresh ? qa : qb);
    if (!s.fresh) {
        R st[5];
#pragma unroll
        for (int k = 0; k < S; ++k) st[k] = ld_stream<2>(a.deps_last + (long long)k * a.N + v);
        expand<R, DIM>(st, s.pd);
        s.norm2 = sqrt(contract<R, DIM>(s.pd, s.pd));                    // :617-618
    }
}

template <typename R>
struct ModeData {
    R q0;           // cached Q0*exp(-D*S) when the voxel is fresh, else Q0
    R st[5];
};

template <typename R, int DIM>
__device__ __forceinline__ void load_mode(const RateArgs<R> &a, long long v, int m, const R *q0src, ModeData<R> &d)
{
    constexpr int S = Comp<DIM>::S;
    d.q0 = ld_stream<1>(q0src + (long long)m * a.N + v);
#pragma unroll
    for (int k = 0; k < S; ++k) d.st[k] = ld_stream<1>(a.deps + ((long long)m * S + k) * a.N + v);
}

// Q and rate of one mode at a voxel; returns Q0*exp(-D*S).  softm: the voxel's softening
// for this mode (s.soft for glass; for crystal-like MRO s.soft times the factor once per
// mode up to this one, the reference scaling Softening[idx] in place in its mode loop)
template <typename R, int DIM, bool XTAL = false>
__device__ __forceinline__ R mode_q(const RateArgs<R> &a, const ModeData<R> &md, const VoxelState<R> &s, R softm,
                                    R &q, double &rate)
{
    constexpr int C = Comp<DIM>::C;
    R d[C];
    expand<R, DIM>(md.st, d);
    const R dwork = contract<R, DIM>(s.sg, d) * (R)a.voxel * (R)1e-18 * (R)STZ_J_IN_EV;   // :585-586
    R qe;
    if (s.fresh) {
        qe = md.q0;
    } else {
        const R dot = contract<R, DIM>(s.pd, d);                                           // :613-614
        const R norm1 = sqrt(contract<R, DIM>(d, d));                                      // :615-616
        R dm = dot / (norm1 + (R)STZ_EPSILON) / (s.norm2 + (R)STZ_EPSILON);                // :620
        dm = ((R)1 + dm) * ((R)1 + dm) / (R)4;                                             // :621
        if constexpr (XTAL)
            if (s.t) dm = dm * s.fac;                                                      // :647-658
        qe = md.q0 * exp(-dm * softm);
    }
    q = qe - dwork / (R)2;                                                                 // :684
    if constexpr (sizeof(R) == 8) rate = a.ratepre * exp(-q / a.kT);                       // evolution.c:125
    else rate = (double)expf((R)a.lnpre - q / (R)a.kT);
    return qe;
}

#pragma clang fp contract(on)

__device__ __forceinline__ int wave_sum_int(int x)
{
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

__device__ __forceinline__ double wave_sum_xor(double x)
{
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x = x + __shfl_xor(x, d, WAVE);
    return x;
}

// inclusive Hillis-Steele scan inside a wavefront; lane 63 equals wave_sum_xor
__device__ __forceinline__ double wave_scan(double x, int lane)
{
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const double y = __shfl_up(x, d, WAVE);
        if (lane >= d) x = y + x;
    }
    return x;
}

// block-wide inclusive scan of NT values: wave scans + sequential wave offsets.
// Its last element equals block_sum() below bit for bit.
__device__ __forceinline__ double block_scan(double x, double *sh)
{
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    x = wave_scan(x, lane);
    if (lane == WAVE - 1) sh[w] = x;
    __syncthreads();
    double off = 0.0;
    for (int i = 0; i < w; ++i) off = off + sh[i];
    __syncthreads();
    return w == 0 ? x : off + x;
}



// ---------------------------------------------------------------------------------
// rate pass over one block of NT voxels (all modes): block rate sums per mode, the
// per-voxel minimum Q and its first argmin, the athermal flag count of the block, the
// Q<0 count, and the Q0*exp(-D*S) cache refresh.  XTAL: the context has crystal-like
// MRO voxels (a.mro set); glass contexts compile without the per-voxel type.  LEAN:
// block rate sums and the Q<0 count only (no cache, no per-voxel minimum, no athermal
// count): the stateless pass of stz_rates_noalloc.
// ---------------------------------------------------------------------------------
template <typename R, int DIM, bool XTAL = false, bool LEAN = false>
__device__ __forceinline__ void rates_block(const RateArgs<R> &a, long long blk)
{
    __shared__ double wsum[STZ_MAXM][NW];
    __shared__ int wcnt[2][NW];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long v = blk * NT + tid;
    const bool valid = v < a.N;
    VoxelState<R> s;
    // modes are loaded DEP ahead of the one being evaluated
    constexpr int DEP = rate_depth<R>();
    ModeData<R> md[DEP + 1];
    if (valid) {
        load_voxel<R, DIM, XTAL, LEAN>(a, v, !LEAN, s);
#pragma unroll
        for (int j = 0; j < DEP; ++j)
            if (j < a.M) load_mode<R, DIM>(a, v, j, s.q0src, md[j]);
    }
    R minq = (R)0;
    int am = 0, nneg = 0;
    R softm = valid ? s.soft : (R)0;
    for (int m = 0; m < a.M; ++m) {
        double rate = 0.0;
        if (valid) {
            if (m + DEP < a.M) load_mode<R, DIM>(a, v, m + DEP, s.q0src, md[DEP]);
            if constexpr (XTAL)
                if (s.t) softm = softm * s.fac;
            R q;
            const R qe = mode_q<R, DIM, XTAL>(a, md[0], s, softm, q, rate);
            if (!LEAN && !s.fresh) a.q0eff[(long long)m * a.N + v] = qe;
            if (m == 0 || minq > q) { minq = q; am = m; }     // first argmin (STZ/evolution.c:56-63)
            nneg += (q < (R)0);
#pragma unroll
            for (int j = 0; j < DEP; ++j) md[j] = md[j + 1];
        }
        const double t = wave_sum_xor(rate);
        if (lane == 0) wsum[m][w] = t;
    }
    const int flag = valid && (minq < (R)0 || fabs((double)minq) < a.thr);   // STZ/evolution.c:65
    const int wn = wave_sum_int(nneg);
    const int wa = __popcll(__ballot(flag));
    if (lane == 0) {
        wcnt[0][w] = wn;
        wcnt[1][w] = wa;
    }
    if (!LEAN && valid) {
        a.minq[v] = minq;
        a.argm[v] = (uint8_t)am;
        // a settled entry holds -S; an S with its sign bit set is not cached (NaN)
        if (!s.fresh || s.settle)
            a.scache[v] = s.settle ? -s.soft : (__builtin_signbit(s.soft) ? (R)__builtin_nan("") : s.soft);
    }
    __syncthreads();
    if (tid < a.M) {
        double t = wsum[tid][0];
        for (int i = 1; i < NW; ++i) t = t + wsum[tid][i];
        a.bsum[(long long)tid * a.nblk + blk] = t;
    }
    if (tid == 0) {
        int tn = 0, ta = 0;
        for (int i = 0; i < NW; ++i) {
            tn += wcnt[0][i];
            ta += wcnt[1][i];
        }
        if (tn) atomicAdd(a.nneg, (unsigned long long)tn);
        if (!LEAN) a.acnt[blk] = ta;
    }
}

// fixed-tree sum of NT consecutive block sums starting at b0 (one per thread)
__device__ __forceinline__ double group_sum(const double *bsum, long long nb, long long b0, double *ws)
{
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long b = b0 + tid;
    double x = b < nb ? bsum[b] : 0.0;
    x = wave_sum_xor(x);
    if (lane == 0) ws[w] = x;
    __syncthreads();
    double t = ws[0];
    for (int i = 1; i < NW; ++i) t = t + ws[i];
    __syncthreads();
    return t;
}

// Running scan over the group sums, chunk by chunk, by one workgroup.  The same code
// computes R and the level-2 prefixes of the selection, so they agree bit for bit.
__device__ double group_total(const double *gsum, long long ng, double *sh)
{
    __shared__ double last;
    constexpr int PRE = 8;            // chunks whose loads are issued before the first scan
    double carry = 0.0;
    for (long long c00 = 0; c00 < ng; c00 += (long long)NT * PRE) {
        double xs[PRE];
#pragma unroll
        for (int i = 0; i < PRE; ++i) {
            const long long g = c00 + (long long)i * NT + threadIdx.x;
            xs[i] = g < ng ? gsum[g] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < PRE; ++i) {
            const long long c0 = c00 + (long long)i * NT;
            if (c0 < ng) {            // uniform over the workgroup
                const double incl = block_scan(xs[i], sh);
                if (threadIdx.x == NT - 1) last = incl;
                __syncthreads();
                carry = (c0 == 0) ? last : carry + last;
                __syncthreads();
            }
        }
    }
    return carry;
}

// The total of chunk c (NT consecutive group sums) exactly as group_total's running scan
// takes it: the last element of the chunk's block scan.  All threads call it.
__device__ __forceinline__ double chunk_total(const double *gsum, long long ng, long long c, double *sh)
{
    const long long g = c * NT + threadIdx.x;
    const double x = g < ng ? gsum[g] : 0.0;
    return block_scan(x, sh);       // thread NT - 1 holds the chunk total
}

// R from the chunk totals: the running carry of group_total, by one thread (bit for bit
// the same value; the chunk totals are computed in parallel beforehand)
__device__ __forceinline__ double chunks_total(const double *csum, long long nch)
{
    double carry = 0.0;
    for (long long c = 0; c < nch; ++c) carry = (c == 0) ? csum[0] : carry + csum[c];
    return carry;
}

// Exclusive offsets of the per-block athermal counts (exact integers), by one
// workgroup: per-thread runs over contiguous chunks, one scan, per-thread writes.
// Returns the total to every thread.
__device__ long long block_offsets(const int *cnt, long long nblk, int *off)
{
    __shared__ long long sh[NT];
    const int tid = threadIdx.x;
    const long long per = (nblk + NT - 1) / NT;
    const long long b0 = (long long)tid * per;
    const long long b1 = b0 + per < nblk ? b0 + per : nblk;
    long long run = 0;
    for (long long b = b0; b < b1; ++b) run += cnt[b];
    sh[tid] = run;
    __syncthreads();
    for (int d = 1; d < NT; d <<= 1) {
        const long long y = tid >= d ? sh[tid - d] : 0;
        __syncthreads();
        sh[tid] += y;
        __syncthreads();
    }
    long long base = sh[tid] - run;
    for (long long b = b0; b < b1; ++b) {
        off[b] = (int)base;
        base += cnt[b];
    }
    const long long total = sh[NT - 1];
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------------
// voxel transform (glass branch of VoxelShearTransform).  The M*S new transformation
// strains are dscale * dnew[i] (dscale = 1 for values drawn already scaled).
// ---------------------------------------------------------------------------------
struct ApplyArgs {
    double *eps0;
    void *deps, *deps_last;
    void *scache;
    double *perm, *temp, *tlast;
    double *tloc;                 // local temperature, heated by each transform
    const void *sigma;            // stress field (storage precision) + soff: the current stress
    double soff[6];
    // crystal-like MRO voxels redraw their M barriers instead of their strains
    void *q0;
    const uint8_t *mro;           // nullptr: glass everywhere
    double mro_q0, rratio;        // xQ0amp, RandomRatio
    const double *unew;           // thermal event: M uniforms drawn for a crystal-like site
    long long N;
    int M;
    double permsoft, tempsoft, permcap, totime;
};

#pragma clang fp contract(off)
// everything but the redraw of the M*S new transformation strains.  Every value is read
// before anything is written (the arrays are distinct and each value is read once), so
// the loads are in flight together rather than one round trip per read-modify-write.
template <typename R, int DIM>
__device__ void transform_voxel_state(const ApplyArgs &a, long long v, int mm)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    const R *deps = (const R *)a.deps;
    R *last = (R *)a.deps_last;
    const R *sg = (const R *)a.sigma;
    R st[5], d[C], sr[C];
    double e0[C];
#pragma unroll
    for (int k = 0; k < S; ++k) st[k] = deps[((long long)mm * S + k) * a.N + v];
#pragma unroll
    for (int c = 0; c < C; ++c) e0[c] = a.eps0[(long long)c * a.N + v];
#pragma unroll
    for (int c = 0; c < C; ++c) sr[c] = sg[(long long)c * a.N + v];
    const double tl = a.tloc[v];
    const double pm = a.perm[v];
    const bool six = a.mro && a.mro[v] == 1;
#pragma unroll
    for (int k = 0; k < S; ++k) last[(long long)k * a.N + v] = st[k];                  // PreDEps0 snapshot :369-374
    expand<R, DIM>(st, d);
    double dd[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dd[c] = (double)d[c];
#pragma unroll
    for (int c = 0; c < C; ++c) a.eps0[(long long)c * a.N + v] = e0[c] + dd[c];      // :377-379
    {
        // local heating (STZ/elastic.c:381-383): |sigma:deps| * 1e9 J/m^3 * (1/Cp/rho)
        double sv[C];
#pragma unroll
        for (int c = 0; c < C; ++c) sv[c] = (double)sr[c] + a.soff[c];
        const double dh = contract<double, DIM>(sv, dd) * 1E9;
        a.tloc[v] = tl + fabs(dh) * STZ_THERMAL_COEFF1;
    }
    const double ds = contract<double, DIM>(dd, dd);                                   // :385-386
    double pv = pm + ds * (six ? STZ_PERM_SOFT_MRO : a.permsoft);                      // :387-390
    if (pv > a.permcap) pv = a.permcap;                                               // :391
    a.perm[v] = pv;
    a.temp[v] = ds * (six ? STZ_TEMP_SOFT_MRO : a.tempsoft);                          // :393-396
    a.tlast[v] = a.totime;                                                            // :397
    ((R *)a.scache)[v] = (R)__builtin_nan("");                                        // Q0*exp(-D*S) must be recomputed
}

template <typename R, int DIM>
__device__ void transform_voxel(const ApplyArgs &a, long long v, int mm, const double *dnew, double dscale)
{
    constexpr int S = Comp<DIM>::S;
    R *deps = (R *)a.deps;
    transform_voxel_state<R, DIM>(a, v, mm);
    if (a.mro && a.mro[v]) {
        // crystal-like MRO: M new barriers xQ0amp*(1 - RandomRatio + RandomRatio*U) from
        // the first M values (uniforms), strains unchanged (STZ/elastic.c:446-506)
        R *q0 = (R *)a.q0;
        for (int m = 0; m < a.M; ++m) q0[(long long)m * a.N + v] = (R)(a.mro_q0 * (1 - a.rratio + a.rratio * (dnew[m])));
        return;
    }
    for (int m = 0; m < a.M; ++m)
#pragma unroll
        for (int k = 0; k < S; ++k)
            deps[((long long)m * S + k) * a.N + v] = (R)(dscale * dnew[m * S + k]);   // :508-512
}

// The redraws of transform_voxel (the crystal-like barriers or the M*S strains) of nv
// voxels vox[j] with draws dsrc + didx[j]*M*S (j < nv, didx[j] < cap), by the NT threads
// of a workgroup: the (voxel, value) pairs are spread over the threads and each thread
// issues 8 loads before its stores, instead of one load-store round trip per value.
template <typename R, int DIM>
__device__ void redraw_wg(const ApplyArgs &a, const long long *vox, const long long *didx, int nv, long long cap,
                          const double *dsrc, double dscale)
{
    constexpr int S = Comp<DIM>::S, U = 8;
    const int MS = a.M * S;
    const int items = nv * MS;
    R *deps = (R *)a.deps;
    R *q0 = (R *)a.q0;
    for (int b0 = 0; b0 < items; b0 += U * NT) {
        double dv[U];
        int xt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = b0 + u * NT + (int)threadIdx.x;
            xt[u] = 0;
            if (it < items && didx[it / MS] < cap) {
                const long long v = vox[it / MS];
                xt[u] = (a.mro && a.mro[v]) ? 2 : 1;
                dv[u] = dsrc[didx[it / MS] * MS + it % MS];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = b0 + u * NT + (int)threadIdx.x;
            if (xt[u] == 0) continue;
            const long long v = vox[it / MS];
            const int i = it % MS;
            if (xt[u] == 2) {   // crystal-like site: M barriers from the first M draws
                if (i < a.M) q0[(long long)i * a.N + v] = (R)(a.mro_q0 * (1 - a.rratio + a.rratio * (dv[u])));
            } else {
                deps[(long long)i * a.N + v] = (R)(dscale * dv[u]);   // :508-512
            }
        }
    }
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------------
// residence-time selection by one workgroup
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int first_true(bool f, int *sh_min)
{
    if (threadIdx.x == 0) *sh_min = NT;
    __syncthreads();
    if (f) atomicMin(sh_min, (int)threadIdx.x);
    __syncthreads();
    const int r = *sh_min;
    __syncthreads();
    return r;
}

// Three-level search for the first event index whose cumulative rate reaches tau
// (STZ/evolution.c:150-224): groups -> blocks -> voxels, each level a wavefront-shuffle
// scan whose last element reproduces the sums of the rate pass bit for bit, the voxel
// level re-evaluating the NT rates of one block with the rate pass's own code.  Writes
// site, mode, Q to res[10..12] and, when dnew is given, transforms the chosen voxel.
// Returns false (res[10] = -1) when tau lies beyond the total.
template <typename R, int DIM, bool XTAL = false>
__device__ bool select_block(const RateArgs<R> &a, const double *gsum, const double *csum, long long ng, double tau,
                             double *res, const ApplyArgs &ap, const double *dnew, double dscale)
{
    __shared__ double sh[NW];
    __shared__ int smin;
    __shared__ double vals[NT];
    __shared__ long long s_cstar;
    __shared__ double s_carry;
    const int tid = threadIdx.x;
    const long long nb = (long long)a.M * a.nblk;
    // level 3: chunks of NT groups, by the running carry of the chunk totals (the values
    // R was summed with); the first chunk whose carry reaches tau
    if (tid == 0) {
        const long long nch = (ng + NT - 1) / NT;
        double carry = 0.0, prev = 0.0;
        long long cs = -1;
        for (long long c = 0; c < nch; ++c) {
            prev = carry;
            carry = (c == 0) ? csum[0] : carry + csum[c];
            if (carry >= tau) {
                cs = c;
                break;
            }
        }
        s_cstar = cs;
        s_carry = prev;
    }
    __syncthreads();
    const long long cstar = s_cstar;
    if (cstar < 0) {                    // tau beyond the total: "failed to locate"
        if (tid == 0) {
            res[10] = -1;
            res[11] = -1;
        }
        return false;
    }
    // level 2: groups of chunk cstar, prefixes carry + block scan as the running scan has them
    long long gstar;
    double base2;
    {
        const long long c0 = cstar * NT;
        const double carry = s_carry;
        const long long g = c0 + tid;
        const double x = g < ng ? gsum[g] : 0.0;
        const double incl = block_scan(x, sh);
        const double P = (c0 == 0) ? incl : carry + incl;
        vals[tid] = P;
        int f = first_true(g < ng && P >= tau, &smin);
        if (f == NT) {                  // rounding put the chunk's last prefix alone at tau
            const long long lastg = (ng - 1 < c0 + NT - 1) ? ng - 1 : c0 + NT - 1;
            f = (int)(lastg - c0);
        }
        gstar = c0 + f;
        base2 = f == 0 ? (c0 == 0 ? 0.0 : carry) : vals[f - 1];
        __syncthreads();
    }
    // level 1: blocks of group gstar
    const long long b = gstar * STZ_GRP + tid;
    const double x = b < nb ? a.bsum[b] : 0.0;
    const double incl = block_scan(x, sh);
    const double P = base2 + incl;
    vals[tid] = P;
    __syncthreads();
    int f = first_true(b < nb && P >= tau, &smin);
    if (f == NT) {
        const long long lastb = (nb - 1 < gstar * STZ_GRP + NT - 1) ? nb - 1 : gstar * STZ_GRP + NT - 1;
        f = (int)(lastb - gstar * STZ_GRP);
    }
    const double base1 = f == 0 ? base2 : vals[f - 1];
    __syncthreads();
    const long long bstar = gstar * STZ_GRP + f;
    // level 0: voxels of block bstar, re-evaluated
    const int m = (int)(bstar / a.nblk);
    const long long blk = bstar % a.nblk;
    const long long v = blk * NT + tid;
    double rate = 0.0;
    R q = (R)0;
    if (v < a.N) {
        VoxelState<R> s;
        ModeData<R> md;
        load_voxel<R, DIM, XTAL>(a, v, false, s);
        load_mode<R, DIM>(a, v, m, s.q0src, md);
        R softm = s.soft;
        if constexpr (XTAL)
            if (s.t)
                for (int j = 0; j <= m; ++j) softm = softm * s.fac;
        (void)mode_q<R, DIM, XTAL>(a, md, s, softm, q, rate);
    }
    const double incl0 = block_scan(rate, sh);
    const double P0 = base1 + incl0;
    int f0 = first_true(v < a.N && P0 >= tau, &smin);
    if (f0 == NT) {
        const long long lastv = (a.N - 1 < blk * NT + NT - 1) ? a.N - 1 : blk * NT + NT - 1;
        f0 = (int)(lastv - blk * NT);
    }
    if (tid == f0) {
        res[10] = (double)(blk * NT + f0);
        res[11] = (double)m;
        res[12] = (double)q;
        if (dnew) transform_voxel_state<R, DIM>(ap, blk * NT + f0, m);   // VoxelShearTransform right away
    }
    if (dnew) {
        // the redraw of all M modes (STZ/elastic.c:508-512), one value per thread, after the
        // chosen mode's strain has been snapshotted above
        constexpr int S = Comp<DIM>::S;
        __syncthreads();
        R *deps = (R *)ap.deps;
        const long long vv = blk * NT + f0;
        if (ap.mro && ap.mro[vv]) {
            // crystal-like MRO site: M new barriers from the event's uniforms (STZ/elastic.c:446-506)
            R *q0 = (R *)ap.q0;
            for (int i = tid; i < a.M; i += NT) {
#pragma clang fp contract(off)
                q0[(long long)i * a.N + vv] = (R)(ap.mro_q0 * (1 - ap.rratio + ap.rratio * (ap.unew[i])));
            }
        } else {
            for (int i = tid; i < a.M * S; i += NT) {
#pragma clang fp contract(off)
                deps[(long long)i * a.N + vv] = (R)(dscale * dnew[i]);
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------
// athermal_plasticity's compaction and transforms (STZ/evolution.c:52-77) for one block
// of NT voxels.  Voxels whose minimum Q over modes is < 0 or within the threshold take
// list positions off[blk] + (rank inside the block), i.e. row-major order; the first
// `cap` of them transform with redraw values dsrc[idx*M*S ...] (the reference draws them
// voxel by voxel in that order).  The block's sum of |DWork| goes to socp[blk].
// ---------------------------------------------------------------------------------
template <typename R>
struct AthArgs {
    ApplyArgs ap;
    const R *minq;
    const uint8_t *argm;
    const R *sigma;
    double soff[6];
    double voxel, thr;
    int *list;
    const int *off;
    const double *dsrc;
    double dscale;
    long long cap;
    long long N;
    double *socp;
};

// the athermal flag of voxel v (STZ/evolution.c:52-56): min_m Q < 0 or |min_m Q| < thr
template <typename R>
__device__ __forceinline__ int athermal_flag(const AthArgs<R> &a, long long v)
{
    const R q = a.minq[v];
    return q < (R)0 || fabs((double)q) < a.thr;
}

// the block's work given every thread's flag f (0 beyond N).  `par` alternates between
// consecutive calls of one workgroup: the per-wave counts are double-buffered, so the
// next block may start while a slower wave still reads this block's counts (every call
// has a barrier after its count writes, so no wave is two blocks ahead).  A block without
// flagged voxels writes +0.0, the value of its tree of zeros.
template <typename R, int DIM>
__device__ void athermal_block_f(const AthArgs<R> &a, long long blk, int f, int par, long long boff)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    __shared__ int ws2[2][NW];
    __shared__ double red[NT];
    int *ws = ws2[par];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long v = blk * NT + tid;
    const unsigned long long bal = __ballot(f);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(bal);
    __syncthreads();
    int loff = 0, tot = 0;
    for (int i = 0; i < NW; ++i) {
        if (i < w) loff += ws[i];
        tot += ws[i];
    }
    if (tot == 0) {                   // uniform over the workgroup
        if (tid == 0) a.socp[blk] = 0.0;
        return;
    }
    __shared__ long long fv[NT], fidx[NT];   // the block's flagged voxels in list order
    double dw = 0.0;
    if (f) {
        const long long idx = boff + loff + before;
        a.list[idx] = (int)v;
        fv[loff + before] = v;
        fidx[loff + before] = idx;
        if (idx < a.cap) {
            const int mm = a.argm[v];
            R sg[C], st[5], d[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                R x = a.sigma[(long long)c * a.N + v];
                if constexpr (sizeof(R) == 4) x = (R)((double)x + a.soff[c]);
                sg[c] = x;
            }
            const R *deps = (const R *)a.ap.deps;
#pragma unroll
            for (int k = 0; k < S; ++k) st[k] = deps[((long long)mm * S + k) * a.N + v];
            expand<R, DIM>(st, d);
            dw = fabs((double)(contract<R, DIM>(sg, d) * (R)a.voxel * (R)1e-18 * (R)STZ_J_IN_EV));   // :66-67
            transform_voxel_state<R, DIM>(a.ap, v, mm);   // reads the chosen strain before the redraw
        }
    }
    __syncthreads();
    redraw_wg<R, DIM>(a.ap, fv, fidx, tot, a.cap, a.dsrc, a.dscale);
    red[tid] = dw;
    __syncthreads();
    for (int s2 = NT / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) red[tid] = red[tid] + red[tid + s2];
        __syncthreads();
    }
    if (tid == 0) a.socp[blk] = red[0];
}

template <typename R, int DIM>
__device__ void athermal_block(const AthArgs<R> &a, long long blk)
{
    const long long v = blk * NT + threadIdx.x;
    athermal_block_f<R, DIM>(a, blk, v < a.N ? athermal_flag(a, v) : 0, 0, (long long)a.off[blk]);
}

// fixed-order sum of the athermal block partials by one workgroup of NTH threads
template <int NTH>
__device__ double ordered_sum(const double *p, long long n)
{
    __shared__ double red[NTH];
    const int tid = threadIdx.x;
    double a = 0.0;
    for (long long i = tid; i < n; i += NTH) a = a + p[i];
    red[tid] = a;
    __syncthreads();
    for (int s2 = NTH / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) red[tid] = red[tid] + red[tid + s2];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// ordered_sum of the athermal block partials over the blocks with flagged voxels only:
// the others' partials are +0.0 (their tree of zeros), which leaves every running sum as
// it is, so the value equals ordered_sum over all n partials bit for bit.  cnt: flagged
// voxels per block; ccnt: per chunk of NT blocks (whole empty chunks are skipped).
template <int NTH>
__device__ double ordered_sum_flagged(const double *p, const int *cnt, const int *ccnt, long long n)
{
    __shared__ double red[NTH];
    const int tid = threadIdx.x;
    double a = 0.0;
    for (long long i = tid; i < n; i += NTH)
        if (ccnt[i / NT] != 0 && cnt[i] != 0) a = a + p[i];
    red[tid] = a;
    __syncthreads();
    for (int s2 = NTH / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) red[tid] = red[tid] + red[tid + s2];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// ordered_sum_flagged<NTV> on a workgroup of NTH threads with the NTV-thread tree's own
// operations: thread t keeps the strided runs of the virtual threads t + j*NTH, the levels
// whose operands belong to one thread add in registers, the others in LDS (bit for bit)
template <int NTV, int NTH>
__device__ double ordered_sum_flagged_n(const double *p, const int *cnt, const int *ccnt, long long n)
{
    static_assert(NTH <= NTV && NTV % NTH == 0, "NTH divides the tree width");
    constexpr int J = NTV / NTH;
    __shared__ double red[NTH];
    const int tid = threadIdx.x;
    double a[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        a[j] = 0.0;
        for (long long i = tid + j * NTH; i < n; i += NTV)
            if (ccnt[i / NT] != 0 && cnt[i] != 0) a[j] = a[j] + p[i];
    }
#pragma unroll
    for (int w = NTV / 2; w >= NTH; w >>= 1)
#pragma unroll
        for (int j = 0; j < w / NTH; ++j) a[j] = a[j] + a[j + w / NTH];
    red[tid] = a[0];
    __syncthreads();
    for (int s2 = NTH / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) red[tid] = red[tid] + red[tid + s2];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// group sums gsum[g] for the STZ_GRP_PER_BLOCK groups g0.. of one workgroup
#define STZ_GRP_PER_BLOCK 4
__device__ __forceinline__ void group_sums(const double *bsum, long long nb, double *gsum, long long ngrp, long long g0,
                                           double *ws)
{
    for (int j = 0; j < STZ_GRP_PER_BLOCK; ++j) {
        const long long g = g0 + j;
        if (g >= ngrp) break;
        const double t = group_sum(bsum, nb, g * STZ_GRP, ws);
        if (threadIdx.x == 0) gsum[g] = t;
    }
}

}  // namespace
