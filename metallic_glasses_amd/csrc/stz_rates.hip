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
    const double *perm, *temp, *tlast;
    R *minq;
    uint8_t *argm;
    double *bsum;
    int *bneg;
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
};

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
    R st[5];
#pragma unroll
    for (int k = 0; k < S; ++k) st[k] = a.deps_last[(long long)k * a.N + v];
    expand<R, DIM>(st, s.pd);
    s.norm2 = sqrt(contract<R, DIM>(s.pd, s.pd));                        // :617-618
}

// Q and rate of mode m at voxel v
template <typename R, int DIM>
__device__ __forceinline__ void mode_q(const RateArgs<R> &a, long long v, int m, const VoxelState<R> &s, R &q, double &rate)
{
    constexpr int S = Comp<DIM>::S, C = Comp<DIM>::C;
    R st[5], d[C];
#pragma unroll
    for (int k = 0; k < S; ++k) st[k] = a.deps[((long long)m * S + k) * a.N + v];
    expand<R, DIM>(st, d);
    const R dwork = contract<R, DIM>(s.sg, d) * (R)a.voxel * (R)1e-18 * (R)STZ_J_IN_EV;   // :585-586
    const R dot = contract<R, DIM>(s.pd, d);                                               // :613-614
    const R norm1 = sqrt(contract<R, DIM>(d, d));                                          // :615-616
    R dm = dot / (norm1 + (R)STZ_EPSILON) / (s.norm2 + (R)STZ_EPSILON);                    // :620
    dm = ((R)1 + dm) * ((R)1 + dm) / (R)4;                                                 // :621
    q = a.q0[(long long)m * a.N + v] * exp(-dm * s.soft) - dwork / (R)2;                   // :684
    if constexpr (sizeof(R) == 8) rate = a.ratepre * exp(-q / a.kT);                       // evolution.c:125
    else rate = (double)expf((R)a.lnpre - q / (R)a.kT);
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
    __shared__ int wneg[NW];
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long blk = blockIdx.x;
    const long long v = blk * NT + tid;
    const bool valid = v < a.N;
    VoxelState<R> s;
    if (valid) load_voxel<R, DIM>(a, v, true, s);
    R minq = (R)0;
    int am = 0, nneg = 0;
    for (int m = 0; m < a.M; ++m) {
        double rate = 0.0;
        if (valid) {
            R q;
            mode_q<R, DIM>(a, v, m, s, q, rate);
            if (m == 0 || minq > q) { minq = q; am = m; }     // first argmin (STZ/evolution.c:56-63)
            nneg += (q < (R)0);
        }
        const double t = wave_sum_xor(rate);
        if (lane == 0) wsum[m][w] = t;
    }
    const int wn = wave_sum_int(nneg);
    if (lane == 0) wneg[w] = wn;
    if (valid) {
        a.minq[v] = minq;
        a.argm[v] = (uint8_t)am;
    }
    __syncthreads();
    if (tid < a.M) {
        double t = wsum[tid][0];
        for (int i = 1; i < NW; ++i) t = t + wsum[tid][i];
        a.bsum[(long long)tid * a.nblk + blk] = t;
    }
    if (tid == 0) {
        int t = 0;
        for (int i = 0; i < NW; ++i) t += wneg[i];
        a.bneg[blk] = t;
    }
}

// fixed-tree sums of STZ_GRP consecutive block sums
__global__ __launch_bounds__(NT) void k_group(const double *__restrict__ bsum, long long nb, double *__restrict__ gsum)
{
    __shared__ double ws[NW];
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
    }
}

// Running scan over the group sums, chunk by chunk; returns the inclusive value of
// element g in *pg when g >= 0.  Shared by k_total and k_select so R is identical.
__device__ double group_total(const double *gsum, long long ng, double *sh)
{
    double carry = 0.0;
    for (long long c0 = 0; c0 < ng; c0 += NT) {
        const long long g = c0 + threadIdx.x;
        const double x = g < ng ? gsum[g] : 0.0;
        const double incl = block_scan(x, sh);
        __shared__ double last;
        if (threadIdx.x == NT - 1) last = incl;
        __syncthreads();
        carry = (c0 == 0) ? last : carry + last;
        __syncthreads();
    }
    return carry;
}

__global__ __launch_bounds__(NT) void k_total(const double *__restrict__ gsum, long long ng, const int *__restrict__ bneg,
                                              long long nblk, double *__restrict__ res)
{
    __shared__ double sh[NW];
    __shared__ long long nsh[NT];
    const double R = group_total(gsum, ng, sh);
    long long n = 0;
    for (long long b = threadIdx.x; b < nblk; b += NT) n += bneg[b];
    nsh[threadIdx.x] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int i = 0; i < NT; ++i) t += nsh[i];
        res[8] = R;
        res[9] = (double)t;
    }
}

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
                                               double *__restrict__ res)
{
    __shared__ double sh[NW];
    __shared__ int smin;
    __shared__ double vals[NT];
    const int tid = threadIdx.x;
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
        if (tid == 0) { res[10] = -1; res[11] = -1; }
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
            load_voxel<R, DIM>(a, v, false, s);
            mode_q<R, DIM>(a, v, m, s, q, rate);
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
        }
    }
}

// ---------------------------------------------------------------------------------
// voxel transform (glass branch of VoxelShearTransform)
// ---------------------------------------------------------------------------------
struct ApplyArgs {
    double *eps0;
    void *deps, *deps_last;
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
    for (int m = 0; m < a.M; ++m)
#pragma unroll
        for (int k = 0; k < S; ++k) deps[((long long)m * S + k) * a.N + v] = (R)dnew[m * S + k];   // :508-512
}
#pragma clang fp contract(on)

template <typename R, int DIM>
__global__ void k_apply_one(ApplyArgs a, long long v, int mm, const double *__restrict__ dnew)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) transform_voxel<R, DIM>(a, v, mm, dnew);
}

// athermal: flags from the per-voxel minimum (STZ/evolution.c:65)
template <typename R>
__global__ __launch_bounds__(NT) void k_acount(const R *__restrict__ minq, long long N, double thr, int *__restrict__ cnt)
{
    __shared__ int ws[NW];
    const long long v = (long long)blockIdx.x * NT + threadIdx.x;
    int f = 0;
    if (v < N) {
        const R q = minq[v];
        f = (q < (R)0 || fabs((double)q) < thr);
    }
    const int t = wave_sum_int(f);
    if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(NT) void k_aoffsets(const int *__restrict__ cnt, long long nblk, int *__restrict__ off,
                                                 double *__restrict__ res)
{
    __shared__ int sh[NT];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long long c0 = 0; c0 < nblk; c0 += NT) {
        const long long b = c0 + threadIdx.x;
        const int x = b < nblk ? cnt[b] : 0;
        sh[threadIdx.x] = x;
        __syncthreads();
        for (int d = 1; d < NT; d <<= 1) {
            const int y = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
            __syncthreads();
            sh[threadIdx.x] += y;
            __syncthreads();
        }
        if (b < nblk) off[b] = (int)(carry + sh[threadIdx.x] - x);
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[NT - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) res[13] = (double)carry;
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
    a.perm = ctx->perm;
    a.temp = ctx->temp;
    a.tlast = ctx->tlast;
    a.minq = (R *)ctx->minq;
    a.argm = ctx->argm;
    a.bsum = ctx->bsum;
    a.bneg = ctx->bneg;
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
int rates_impl(stz_ctx *ctx)
{
    RateArgs<R> a = rate_args<R>(ctx);
    tmark(ctx, 8);
    hipLaunchKernelGGL((k_rates<R, DIM>), dim3((unsigned)ctx->nblk), dim3(NT), 0, ctx->stream, a);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 9);
    const long long nb = (long long)ctx->M * ctx->nblk;
    hipLaunchKernelGGL(k_group, dim3((unsigned)ctx->ngrp), dim3(NT), 0, ctx->stream, ctx->bsum, nb, ctx->gsum);
    STZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_total, dim3(1), dim3(NT), 0, ctx->stream, ctx->gsum, ctx->ngrp, ctx->bneg, ctx->nblk, ctx->res);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 10);
    return STZ_OK;
}

template <typename R, int DIM>
int select_impl(stz_ctx *ctx, double tau)
{
    RateArgs<R> a = rate_args<R>(ctx);
    a.pend_comp = -1;
    tmark(ctx, 11);
    hipLaunchKernelGGL((k_select<R, DIM>), dim3(1), dim3(NT), 0, ctx->stream, a, ctx->gsum, ctx->ngrp, tau, ctx->res);
    STZ_CHECK_LAUNCH();
    tmark(ctx, 12);
    return STZ_OK;
}

}  // namespace

int stz_rates_launch(stz_ctx *ctx)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? rates_impl<double, 2>(ctx) : rates_impl<double, 3>(ctx);
    return ctx->dim == 2 ? rates_impl<float, 2>(ctx) : rates_impl<float, 3>(ctx);
}

int stz_select_launch(stz_ctx *ctx, double target, double)
{
    if (ctx->prec == STZ_FP64) return ctx->dim == 2 ? select_impl<double, 2>(ctx, target) : select_impl<double, 3>(ctx, target);
    return ctx->dim == 2 ? select_impl<float, 2>(ctx, target) : select_impl<float, 3>(ctx, target);
}

int stz_apply_launch(stz_ctx *ctx, int site, int mode)
{
    ApplyArgs a = apply_args(ctx);
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<double, 2>), dim3(1), dim3(64), 0, ctx->stream, a, (long long)site, mode, ctx->dnew);
        else hipLaunchKernelGGL((k_apply_one<double, 3>), dim3(1), dim3(64), 0, ctx->stream, a, (long long)site, mode, ctx->dnew);
    } else {
        if (ctx->dim == 2) hipLaunchKernelGGL((k_apply_one<float, 2>), dim3(1), dim3(64), 0, ctx->stream, a, (long long)site, mode, ctx->dnew);
        else hipLaunchKernelGGL((k_apply_one<float, 3>), dim3(1), dim3(64), 0, ctx->stream, a, (long long)site, mode, ctx->dnew);
    }
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_athermal_launch(stz_ctx *ctx)
{
    const double thr = ctx->k.thr;
    const dim3 g((unsigned)ctx->nblk);
    if (ctx->prec == STZ_FP64)
        hipLaunchKernelGGL((k_acount<double>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->minq, ctx->N, thr, ctx->acnt);
    else
        hipLaunchKernelGGL((k_acount<float>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->minq, ctx->N, thr, ctx->acnt);
    STZ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_aoffsets, dim3(1), dim3(NT), 0, ctx->stream, ctx->acnt, ctx->nblk, ctx->aoff, ctx->res);
    STZ_CHECK_LAUNCH();
    if (ctx->prec == STZ_FP64)
        hipLaunchKernelGGL((k_acompact<double>), g, dim3(NT), 0, ctx->stream, (const double *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    else
        hipLaunchKernelGGL((k_acompact<float>), g, dim3(NT), 0, ctx->stream, (const float *)ctx->minq, ctx->N, thr, ctx->aoff, ctx->alist);
    STZ_CHECK_LAUNCH();
    return STZ_OK;
}

int stz_athermal_apply_launch(stz_ctx *ctx, int n)
{
    if (n <= 0) return STZ_OK;
    ApplyArgs a = apply_args(ctx);
    const dim3 g((unsigned)((n + NT - 1) / NT));
    double *soff = ctx->res + 16;   // device copy of the mixed-precision offsets
    if (ctx->prec == STZ_FP64) {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<double, 2>), g, dim3(NT), 0, ctx->stream, a, ctx->alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<double, 3>), g, dim3(NT), 0, ctx->stream, a, ctx->alist, n, ctx->argm, (const double *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
    } else {
        if (ctx->dim == 2)
            hipLaunchKernelGGL((k_aapply<float, 2>), g, dim3(NT), 0, ctx->stream, a, ctx->alist, n, ctx->argm, (const float *)ctx->sigma,
                               ctx->pend_delta, ctx->pend_comp, soff, ctx->k.voxel, ctx->dnew, ctx->socp);
        else
            hipLaunchKernelGGL((k_aapply<float, 3>), g, dim3(NT), 0, ctx->stream, a, ctx->alist, n, ctx->argm, (const float *)ctx->sigma,
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
