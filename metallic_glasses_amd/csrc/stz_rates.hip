// stz_rates.hip -- activation rates, residence-time selection, athermal scan and the
// eigenstrain update.
//
// k_rates   one pass over all voxels: SofteningFactor (STZ/init.c:711-739), ActEnergy
//           (STZ/init.c:561-690), the Q<0 count (STZ/bmg.c:147-153), the KMC rates
//           (STZ/evolution.c:120-128) -- never materialised -- reduced to FP64 sums per
//           (mode, block of 256 voxels) in a fixed tree, plus the per-voxel minimum Q
//           and its first argmin for the athermal scan (STZ/evolution.c:55-64).
// k_group   fixed-tree sums of 256 block sums.
// k_total   fixed-order running scan over the group sums -> R.
// k_select  three-level search for the first event index whose cumulative rate reaches
//           dice*R (STZ/evolution.c:150-224): groups -> blocks -> voxels, each level a
//           wavefront-shuffle scan whose last element reproduces the sums above bit for
//           bit, the voxel level re-evaluating the 256 rates of one block with the same
//           code as k_rates.
// k_apply   VoxelShearTransform (STZ/elastic.c:361-520) for one or many voxels.
#include "stz_internal.h"

namespace {

constexpr int NT = 256;
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
    bool fresh;     // the cached Q0*exp(-D*S) of this voxel is valid for this step's S
};

// Softening S (STZ/init.c:731) and stress of one voxel.  Q0*exp(-D*S) (STZ/init.c:684)
// changes only when S changes or the voxel transforms, so it is cached per (mode,
// voxel) with the S it was computed for; S equals Perm bit for bit except shortly after
// a transform, so almost every voxel takes the cached value, which is the bit-identical
// result of the same expression.
template <typename R, int DIM>
__device__ __forceinline__ void load_voxel(const RateArgs<R> &a, long long v, bool apply_pending, VoxelState<R> &s)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    double soft = 0.0;
    if (a.softening) soft = a.perm[v] + a.temp[v] * exp((a.tlast[v] - a.totime) / a.tTemp);   // :731
    s.soft = (R)soft;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        R x = a.sigma[(long long)c * a.N + v];
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
    s.fresh = (a.scache[v] == s.soft);       // false for the NaN of an invalidated entry
    if (!s.fresh) {
        R st[5];
#pragma unroll
        for (int k = 0; k < S; ++k) st[k] = a.deps_last[(long long)k * a.N + v];
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
__device__ __forceinline__ void load_mode(const RateArgs<R> &a, long long v, int m, bool fresh, ModeData<R> &d)
{
    constexpr int S = Comp<DIM>::S;
    d.q0 = fresh ? a.q0eff[(long long)m * a.N + v] : a.q0[(long long)m * a.N + v];
#pragma unroll
    for (int k = 0; k < S; ++k) d.st[k] = a.deps[((long long)m * S + k) * a.N + v];
}

// Q and rate of one mode at a voxel; returns Q0*exp(-D*S)
template <typename R, int DIM>
__device__ __forceinline__ R mode_q(const RateArgs<R> &a, const ModeData<R> &md, const VoxelState<R> &s, R &q,
                                    double &rate)
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
        qe = md.q0 * exp(-dm * s.soft);
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
// k_rates
// ---------------------------------------------------------------------------------
template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_rates(RateArgs<R> a)
{
    __shared__ double wsum[STZ_MAXM][NW];
    __shared__ int wcnt[2][NW];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long blk = blockIdx.x;
    const long long v = blk * NT + tid;
    const bool valid = v < a.N;
    VoxelState<R> s;
    ModeData<R> cur, nxt;
    if (valid) {
        load_voxel<R, DIM>(a, v, true, s);
        load_mode<R, DIM>(a, v, 0, s.fresh, cur);
    }
    R minq = (R)0;
    int am = 0, nneg = 0;
    for (int m = 0; m < a.M; ++m) {
        double rate = 0.0;
        if (valid) {
            // the next mode's loads are in flight while this mode is evaluated
            if (m + 1 < a.M) load_mode<R, DIM>(a, v, m + 1, s.fresh, nxt);
            R q;
            const R qe = mode_q<R, DIM>(a, cur, s, q, rate);
            if (!s.fresh) a.q0eff[(long long)m * a.N + v] = qe;
            if (m == 0 || minq > q) { minq = q; am = m; }     // first argmin (STZ/evolution.c:56-63)
            nneg += (q < (R)0);
            cur = nxt;
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
    if (valid) {
        a.minq[v] = minq;
        a.argm[v] = (uint8_t)am;
        if (!s.fresh) a.scache[v] = s.soft;
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
        a.acnt[blk] = ta;
    }
}

// Step control block in device memory (doubles unless noted):
//   res[0..5]   <sigma> of the last solve        res[8]  R           res[9]  #(Q<0)
//   res[10..12] selected site, mode, Q           res[13] #athermal   res[15] NaN flag (int)
//   res[16..21] mixed-precision stress offsets   res[24] Q<0 counter (u64)
//   res[25]     group-sum arrival counter (u32)  res[26] step kind (int)
//   res[27]     selection target dice*R          res[28] dice (speculative)

// Running scan over the group sums, chunk by chunk, by one workgroup.  The same code
// computes R here and the level-2 prefixes in k_select, so they agree bit for bit.
__device__ double group_total(const double *gsum, long long ng, double *sh)
{
    __shared__ double last;
    double carry = 0.0;
    for (long long c0 = 0; c0 < ng; c0 += NT) {
        const long long g = c0 + threadIdx.x;
        const double x = g < ng ? gsum[g] : 0.0;
        const double incl = block_scan(x, sh);
        if (threadIdx.x == NT - 1) last = incl;
        __syncthreads();
        carry = (c0 == 0) ? last : carry + last;
        __syncthreads();
    }
    return carry;
}

// Fixed-tree sums of STZ_GRP consecutive block sums; the last workgroup to finish
// computes R, takes the Q<0 count and, when `decide` is set, the step kind
// (STZ/bmg.c:161, STZ/evolution.c:134-136) and the selection target.
__global__ __launch_bounds__(NT) void k_group(const double *__restrict__ bsum, long long nb, double *__restrict__ gsum,
                                              long long ng, double *__restrict__ res, int decide, double dt,
                                              const double *dicep)
{
    __shared__ double ws[NW];
    __shared__ int is_last;
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long b = (long long)blockIdx.x * STZ_GRP + tid;
    double x = b < nb ? bsum[b] : 0.0;
    x = wave_sum_xor(x);
    if (lane == 0) ws[w] = x;
    __syncthreads();
    if (tid == 0) {
        double t = ws[0];
        for (int i = 1; i < NW; ++i) t = t + ws[i];
        gsum[blockIdx.x] = t;
        __threadfence();
        unsigned int *arrive = reinterpret_cast<unsigned int *>(res + RES_ARRIVE);
        is_last = (atomicAdd(arrive, 1u) == (unsigned int)(ng - 1));
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    const double R = group_total(gsum, ng, ws);
    if (tid == 0) {
        unsigned long long *cnt = reinterpret_cast<unsigned long long *>(res + RES_NNEG);
        const unsigned long long nn = atomicAdd(cnt, 0ull);
        atomicExch(cnt, 0ull);
        atomicExch(reinterpret_cast<unsigned int *>(res + RES_ARRIVE), 0u);
        res[8] = R;
        res[9] = (double)nn;
        if (decide) {
            int kind;
            if (nn > 0) kind = KIND_ATHERMAL;
            else kind = (1 / R < dt) ? KIND_THERMAL : KIND_ELASTIC;
            const double dice = *dicep;
            if (kind == KIND_THERMAL && !(dice > 0.0)) kind = KIND_FAIL;
            reinterpret_cast<int *>(res + RES_KIND)[0] = kind;
            res[RES_TAU] = dice * R;
            reinterpret_cast<int *>(res + 15)[0] = 0;     // NaN flag of the coming solve
            res[13] = 0.0;
            res[14] = 0.0;                                 // athermal sum |DWork|
        }
    }
}

// ---------------------------------------------------------------------------------
// voxel transform (glass branch of VoxelShearTransform)
// ---------------------------------------------------------------------------------
struct ApplyArgs {
    double *eps0;
    void *deps, *deps_last;
    void *scache;
    double *perm, *temp, *tlast;
    long long N;
    int M;
    double permsoft, tempsoft, permcap, totime;
};

#pragma clang fp contract(off)
template <typename R, int DIM>
__device__ void transform_voxel(const ApplyArgs &a, long long v, int mm, const double *dnew)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    R *deps = (R *)a.deps;
    R *last = (R *)a.deps_last;
    R st[5], d[C];
#pragma unroll
    for (int k = 0; k < S; ++k) st[k] = deps[((long long)mm * S + k) * a.N + v];
#pragma unroll
    for (int k = 0; k < S; ++k) last[(long long)k * a.N + v] = st[k];                  // PreDEps0 snapshot :369-374
    expand<R, DIM>(st, d);
    double dd[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dd[c] = (double)d[c];
#pragma unroll
    for (int c = 0; c < C; ++c) a.eps0[(long long)c * a.N + v] += dd[c];             // :377-379
    const double ds = contract<double, DIM>(dd, dd);                                   // :385-386
    double pv = a.perm[v] + ds * a.permsoft;                                          // :390
    if (pv > a.permcap) pv = a.permcap;                                               // :391
    a.perm[v] = pv;
    a.temp[v] = ds * a.tempsoft;                                                      // :396
    a.tlast[v] = a.totime;                                                            // :397
    ((R *)a.scache)[v] = (R)__builtin_nan("");                                        // Q0*exp(-D*S) must be recomputed
    for (int m = 0; m < a.M; ++m)
#pragma unroll
        for (int k = 0; k < S; ++k) deps[((long long)m * S + k) * a.N + v] = (R)dnew[m * S + k];   // :508-512
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------------
// k_select: one workgroup
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

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_select(RateArgs<R> a, const double *__restrict__ gsum, long long ng, double tau,
                                               double *__restrict__ res, int gated, ApplyArgs ap,
                                               const double *__restrict__ dnew)
{
    __shared__ double sh[NW];
    __shared__ int smin;
    __shared__ double vals[NT];
    const int tid = threadIdx.x;
    if (gated) {
        if (reinterpret_cast<const int *>(res + RES_KIND)[0] != KIND_THERMAL) return;
        tau = res[RES_TAU];
    }
    const long long nb = (long long)a.M * a.nblk;
    // level 2: groups
    double carry = 0.0;
    long long gstar = -1;
    double base2 = 0.0;
    for (long long c0 = 0; c0 < ng && gstar < 0; c0 += NT) {
        const long long g = c0 + tid;
        const double x = g < ng ? gsum[g] : 0.0;
        const double incl = block_scan(x, sh);
        const double P = (c0 == 0) ? incl : carry + incl;
        vals[tid] = P;
        const int f = first_true(g < ng && P >= tau, &smin);
        if (f < NT) {
            gstar = c0 + f;
            base2 = f == 0 ? (c0 == 0 ? 0.0 : carry) : vals[f - 1];
        } else {
            carry = vals[NT - 1];
        }
        __syncthreads();
    }
    if (gstar < 0) {                    // tau beyond the total: "failed to locate"
        if (tid == 0) {
            res[10] = -1;
            res[11] = -1;
            reinterpret_cast<int *>(res + RES_KIND)[0] = KIND_FAIL;
        }
        return;
    }
    // level 1: blocks of group gstar
    const long long b = gstar * STZ_GRP + tid;
    {
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
            load_voxel<R, DIM>(a, v, false, s);
            load_mode<R, DIM>(a, v, m, s.fresh, md);
            (void)mode_q<R, DIM>(a, md, s, q, rate);
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
            // fused step: VoxelShearTransform of the chosen (mode, site) right away
            if (dnew) transform_voxel<R, DIM>(ap, blk * NT + f0, m, dnew);
        }
    }
}

// ---------------------------------------------------------------------------------
// k_athermal: athermal_plasticity's scan and transforms (STZ/evolution.c:52-77) in one
// pass.  Voxels whose minimum Q over modes is < 0 or within the threshold are
// compacted in row-major order with a decoupled look-back over per-block counts (exact
// integers), and each of the first `cap` of them is transformed with its slice of the
// host-drawn redraw values, whose order is the reference's (voxel by voxel, row-major).
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
    unsigned long long *status;   // per block: epoch << 34 | state << 32 | count
    unsigned int *ticket;
    const double *dnew;           // cap voxels x M*S redraw values
    long long cap;
    unsigned long long epoch;
    long long N, nblk;
    double *res;
    int gated;
};

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_athermal(AthArgs<R> a)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    if (a.gated && reinterpret_cast<const int *>(a.res + RES_KIND)[0] != KIND_ATHERMAL) return;
    __shared__ int s_blk;
    __shared__ int ws[NW];
    __shared__ long long s_excl;
    __shared__ double red[NT];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    // dynamic block order: every block with a smaller ticket has already started
    if (tid == 0) s_blk = (int)atomicAdd(a.ticket, 1u);
    __syncthreads();
    const long long blk = s_blk;
    const long long v = blk * NT + tid;
    int f = 0;
    if (v < a.N) {
        const R q = a.minq[v];
        f = (q < (R)0 || fabs((double)q) < a.thr);
    }
    const unsigned long long bal = __ballot(f);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(bal);
    __syncthreads();
    int loff = 0, agg = 0;
    for (int i = 0; i < NW; ++i) {
        if (i < w) loff += ws[i];
        agg += ws[i];
    }
    if (tid == 0) {
        const unsigned long long ep = a.epoch << 34;
        long long excl = 0;
        if (blk > 0) {
            __hip_atomic_store(&a.status[blk], ep | (1ull << 32) | (unsigned long long)agg, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            long long j = blk - 1;
            for (;;) {
                const unsigned long long st = __hip_atomic_load(&a.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((st >> 34) != a.epoch) continue;          // not published yet in this launch
                excl += (long long)(st & 0xffffffffull);
                if (((st >> 32) & 3ull) == 2ull) break;       // inclusive prefix reached
                --j;
            }
        }
        __hip_atomic_store(&a.status[blk], ep | (2ull << 32) | (unsigned long long)(excl + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        s_excl = excl;
        if (blk == a.nblk - 1) {             // the last ticket: the total is known, reset the ticket
            a.res[13] = (double)(excl + agg);
            if (excl + agg > a.cap) reinterpret_cast<int *>(a.res + RES_KIND)[0] = KIND_ATH_OVERFLOW;
            atomicExch(a.ticket, 0u);
        }
    }
    __syncthreads();
    double dw = 0.0;
    if (f) {
        const long long idx = s_excl + loff + before;
        a.list[idx] = (int)v;
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
            transform_voxel<R, DIM>(a.ap, v, mm, a.dnew + idx * (long long)a.ap.M * S);
        }
    }
    red[tid] = dw;
    __syncthreads();
    for (int s2 = NT / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) red[tid] += red[tid + s2];
        __syncthreads();
    }
    if (tid == 0 && red[0] != 0.0) atomicAdd(a.res + 14, red[0]);
}

template <typename R, int DIM>
__global__ void k_apply_one(ApplyArgs a, long long v, int mm, const double *__restrict__ dnew, const double *res)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (res) {      // fused step: the event comes from k_select, only on a thermal step
        if (reinterpret_cast<const int *>(res + RES_KIND)[0] != KIND_THERMAL) return;
        v = (long long)res[10];
        mm = (int)res[11];
    }
    transform_voxel<R, DIM>(a, v, mm, dnew);
}

// athermal: exclusive offsets of the per-block counts written by k_rates
// (three phases: per-thread runs, one block scan, per-thread writes)
constexpr int NT_OFF = 1024;
__global__ __launch_bounds__(NT_OFF) void k_aoffsets(const int *__restrict__ cnt, long long nblk, int *__restrict__ off,
                                                     double *__restrict__ res)
{
    __shared__ long long sh[NT_OFF];
    const int tid = threadIdx.x;
    const long long per = (nblk + NT_OFF - 1) / NT_OFF;
    const long long b0 = (long long)tid * per;
    const long long b1 = b0 + per < nblk ? b0 + per : nblk;
    long long run = 0;
    for (long long b = b0; b < b1; ++b) run += cnt[b];
    sh[tid] = run;
    __syncthreads();
    for (int d = 1; d < NT_OFF; d <<= 1) {
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
    if (tid == NT_OFF - 1) res[13] = (double)sh[NT_OFF - 1];
}

template <typename R>
__global__ __launch_bounds__(NT) void k_acompact(const R *__restrict__ minq, long long N, double thr,
                                                 const int *__restrict__ off, int *__restrict__ list)
{
    __shared__ int ws[NW];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long v = (long long)blockIdx.x * NT + tid;
    int f = 0;
    if (v < N) {
        const R q = minq[v];
        f = (q < (R)0 || fabs((double)q) < thr);
    }
    const unsigned long long bal = __ballot(f);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(bal);
    __syncthreads();
    int base = off[blockIdx.x];
    for (int i = 0; i < w; ++i) base += ws[i];
    if (f) list[base + before] = (int)v;
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_aapply(ApplyArgs a, const int *__restrict__ list, int n, const uint8_t *__restrict__ argm,
                                               const R *__restrict__ sigma, double pend, int pcomp, const double *soff,
                                               double voxel, const double *__restrict__ dnew, double *__restrict__ socp)
{
    constexpr int C = Comp<DIM>::C, S = Comp<DIM>::S;
    __shared__ double red[NT];
    const int i = blockIdx.x * NT + threadIdx.x;
    double dw = 0.0;
    if (i < n) {
        const long long v = list[i];
        const int mm = argm[v];
        // DWork of the chosen mode with the current stress (STZ/evolution.c:66-67)
        R sg[C], st[5], d[C];
        for (int c = 0; c < C; ++c) {
            R x = sigma[(long long)c * a.N + v];
            if constexpr (sizeof(R) == 8) { if (c == pcomp) x = x + (R)pend; }
            else x = (R)((double)x + soff[c]);
            sg[c] = x;
        }
        const R *deps = (const R *)a.deps;
        for (int k = 0; k < S; ++k) st[k] = deps[((long long)mm * S + k) * a.N + v];
        expand<R, DIM>(st, d);
        dw = fabs((double)(contract<R, DIM>(sg, d) * (R)voxel * (R)1e-18 * (R)STZ_J_IN_EV));
        transform_voxel<R, DIM>(a, v, mm, dnew + (long long)i * a.M * S);
    }
    red[threadIdx.x] = dw;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) socp[blockIdx.x] = red[0];
}

// elastic energy partials (ElasticEnergy, STZ/elastic.c:533-547)
#pragma clang fp contract(off)
template <typename R, int DIM>
__global__ __launch_bounds__(NT) void k_energy(const R *__restrict__ sigma, long long N, double lam, double mu, int pcomp,
                                               double pend, const double *soff, double *__restrict__ part)
{
    constexpr int C = Comp<DIM>::C;
    __shared__ double red[NT];
    const long long v = (long long)blockIdx.x * NT + threadIdx.x;
    double e = 0.0;
    if (v < N) {
        double s[C], el[C];
        for (int c = 0; c < C; ++c) {
            double x;
            if constexpr (sizeof(R) == 8) { x = sigma[(long long)c * N + v]; if (c == pcomp) x = x + pend; }
            else x = (double)sigma[(long long)c * N + v] + soff[c];
            s[c] = x;
        }
        if constexpr (DIM == 2) {
            double tra = -lam / 2 / mu / (2 * lam + 2 * mu) * (s[0] + s[1]);
            el[0] = tra + s[0] / 2 / mu;
            el[1] = tra + s[1] / 2 / mu;
            el[2] = s[2] / 2 / mu;
        } else {
            double tra = -lam / 2 / mu / (3 * lam + 2 * mu) * ((s[0] + s[1]) + s[2]);
            for (int c = 0; c < 3; ++c) el[c] = tra + s[c] / 2 / mu;
            for (int c = 3; c < 6; ++c) el[c] = s[c] / 2 / mu;
        }
        e = 0.5 * contract<double, DIM>(s, el);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
#pragma clang fp contract(on)

template <typename R>
RateArgs<R> rate_args(stz_ctx *ctx)
{
    RateArgs<R> a;
    a.sigma = (R *)ctx->sigma;
    a.q0 = (const R *)ctx->q0;
    a.deps = (const R *)ctx->deps;
    a.deps_last = (const R *)ctx->deps_last;
    a.q0eff = (R *)ctx->q0eff;
    a.scache = (R *)ctx->scache;
    a.perm = ctx->perm;
    a.temp = ctx->temp;
    a.tlast = ctx->tlast;
    a.minq = (R *)ctx->minq;
    a.argm = ctx->argm;
    a.bsum = ctx->bsum;
    a.acnt = ctx->acnt;
    a.nneg = reinterpret_cast<unsigned long long *>(ctx->res + RES_NNEG);
    a.thr = ctx->k.thr;
    a.N = ctx->N;
    a.nblk = ctx->nblk;
    a.M = ctx->M;
    a.softening = ctx->p.softening == 1;
    a.pend_comp = ctx->pend_comp;
    a.pend_delta = ctx->pend_delta;
    for (int c = 0; c < 6; ++c) a.soff[c] = ctx->soff[c];
    a.totime = ctx->totime;
    a.tTemp = ctx->k.tTemp;
    a.kT = ctx->k.kT;
    a.ratepre = ctx->k.ratepre;
    a.voxel = ctx->k.voxel;
    a.lnpre = std::log(ctx->k.ratepre);
    return a;
}

ApplyArgs apply_args(stz_ctx *ctx)
{
    ApplyArgs a;
    a.eps0 = ctx->eps0;
    a.deps = ctx->deps;
    a.deps_last = ctx->deps_last;
    a.scache = ctx->scache;
    a.perm = ctx->perm;
    a.temp = ctx->temp;
    a.tlast = ctx->tlast;
    a.N = ctx->N;
    a.M = ctx->M;
    a.permsoft = ctx->k.permsoft;
    a.tempsoft = ctx->k.tempsoft;
    a.permcap = ctx->k.permcap;
    a.totime = ctx->totime;
    return a;
}

template <typename R, int DIM>
int rates_impl(stz_ctx *ctx, int decide)
{
    RateArgs<R> a = rate_args<R>(ctx);
    tmark(ctx, 8);
    hipLaunchKernelGGL((k_rates<R, DIM>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 9);
    const long long nb = (long long)ctx->M * ctx->nblk;
    hipLaunchKernelGGL(k_group, dim3((unsigned)ctx->ngrp), dim3(NT), 0, ctx->stream, ctx->bsum, nb, ctx->gsum, ctx->ngrp,
                       ctx->res, decide, ctx->k.dt, ctx->dspec_dev);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 10);
    return STZ_OK;
}

template <typename R, int DIM>
int select_impl(stz_ctx *ctx, double tau, int gated)
{
    RateArgs<R> a = rate_args<R>(ctx);
    a.pend_comp = -1;
    tmark(ctx, 11);
    const double *dnew = gated ? ctx->dspec_dev + 1 : nullptr;
    hipLaunchKernelGGL((k_select<R, DIM>), dim3(1), dim3(NT), 0, ctx->stream, a, ctx->gsum, ctx->ngrp, tau, ctx->res,
                       gated, apply_args(ctx), dnew);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 12);
    return STZ_OK;
}

}  // namespace

int stz_rates_launch(stz_ctx *ctx, int decide)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? rates_impl<double, 2>(ctx, decide) : rates_impl<double, 3>(ctx, decide);
    return ctx->dim == 2 ? rates_impl<float, 2>(ctx, decide) : rates_impl<float, 3>(ctx, decide);
}

int stz_select_launch(stz_ctx *ctx, double target, int gated)
{
    if (ctx->prec == STZ_FP64)
        return ctx->dim == 2 ? select_impl<double, 2>(ctx, target, gated) : select_impl<double, 3>(ctx, target, gated);
    return ctx->dim == 2 ? select_impl<float, 2>(ctx, target, gated) : select_impl<float, 3>(ctx, target, gated);
}

int stz_apply_launch(stz_ctx *ctx, int site, int mode, int gated)
{
    ApplyArgs a = apply_args(ctx);
    const double *res = gated ? ctx->res : nullptr;
    const double *dnew = gated ? ctx->dspec_dev + 1 : ctx->dnew;
    const long long v = site;
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<double, 2>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
        else hipLaunchKernelGGL((k_apply_one<double, 3>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
    } else {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<float, 2>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
        else hipLaunchKernelGGL((k_apply_one<float, 3>), dim3(1), dim3(64), 0, ctx->stream, a, v, mode, dnew, res);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_athermal_launch(stz_ctx *ctx)
{
    const double thr = ctx->k.thr;
    const dim3 g((unsigned)ctx->nblk);
    tmark(ctx, 15);
    hipLaunchKernelGGL(k_aoffsets, dim3(1), dim3(NT_OFF), 0, ctx->stream, ctx->acnt, ctx->nblk, ctx->aoff, ctx->res);
    STZ_CHECK_LAUNCH();
    if (ctx->prec == STZ_FP64)
        hipLaunchKernelGGL((k_acompact<double>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    else
        hipLaunchKernelGGL((k_acompact<float>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 16);
    return STZ_OK;
}

template <typename R, int DIM>
int athermal_fused_impl(stz_ctx *ctx, int gated)
{
    AthArgs<R> a;
    a.ap = apply_args(ctx);
    a.minq = (const R *)ctx->minq;
    a.argm = ctx->argm;
    a.sigma = (const R *)ctx->sigma;
    for (int c = 0; c < 6; ++c) a.soff[c] = ctx->soff[c];
    a.voxel = ctx->k.voxel;
    a.thr = ctx->k.thr;
    a.list = ctx->alist;
    a.status = ctx->astatus;
    a.ticket = reinterpret_cast<unsigned int *>(ctx->res + RES_TICKET);
    a.dnew = ctx->dath_dev;
    a.cap = ctx->acap;
    a.epoch = ++ctx->aepoch;
    a.N = ctx->N;
    a.nblk = ctx->nblk;
    a.res = ctx->res;
    a.gated = gated;
    tmark(ctx, 15);
    hipLaunchKernelGGL((k_athermal<R, DIM>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 16);
    return STZ_OK;
}

int stz_athermal_fused_launch(stz_ctx *ctx, int gated)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? athermal_fused_impl<double, 2>(ctx, gated) : athermal_fused_impl<double, 3>(ctx, gated);
    return ctx->dim == 2 ? athermal_fused_impl<float, 2>(ctx, gated) : athermal_fused_impl<float, 3>(ctx, gated);
}

int stz_athermal_apply_launch(stz_ctx *ctx, int n) { return stz_athermal_apply_range(ctx, 0, n); }

int stz_athermal_apply_range(stz_ctx *ctx, long long first, long long count)
{
    const int n = (int)count;
    if (n <= 0) return STZ_OK;
    ApplyArgs a = apply_args(ctx);
    const int *alist = ctx->alist + first;
    const dim3 g((unsigned)((n + NT - 1) / NT));
    double *soff = ctx->res + 16;   // device copy of the mixed-precision offsets
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<double, 2>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<double, 3>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<float, 2>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const float *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<float, 3>), g, dim3(NT), 0, ctx->stream, a, alist, n, ctx->argm, (const float *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_energy_launch(stz_ctx *ctx)
{
    const dim3 g((unsigned)ctx->nblk);
    double *soff = ctx->res + 16;
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_energy<double, 2>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
        else
            hipLaunchKernelGGL((k_energy<double, 3>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_energy<float, 2>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
        else
            hipLaunchKernelGGL((k_energy<float, 3>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->sigma, ctx->N, ctx->k.lambda, ctx->k.mu,
                               ctx->pend_comp, ctx->pend_delta, soff, ctx->socp);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}
