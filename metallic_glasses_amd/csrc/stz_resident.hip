// stz_resident.hip -- the device-resident kMC step: one iteration of the reference's main
// loop (STZ/bmg.c:138-411: SofteningFactor + ActEnergy, the Q<0 test, athermal_plasticity
// or KMC with its loading and clock update, the stress solves) with no host round trip,
// for a batch of replicas of one grid shape at a time.
//
// Per step and shape class the stream holds its launches, every one batched over the
// replicas of the class (grid y/z = replica) and gated on the device:
//   kb_rates      rate pass (shared code with k_rates)
//   kb_group      group sums of the block rate sums
//   kb_chunks     chunk totals of the group sums, athermal counts of the block chunks
//   kb_decide     one workgroup per replica: R, the Q<0 count, the step kind, the
//                 residence-time selection and the transform (thermal), or the
//                 avalanche list offsets (athermal)
//   kb_athermal   athermal compaction + transforms
//   5 solve kernels (stz_solve.hip, batched), or kb_local (the event-local update)
//   kb_finish     <sigma>, the |DWork| sum, clock, loading (STZ/evolution.c:262-358),
//                 athermal correction (STZ/evolution.c:91-97), print-step and
//                 termination tests (STZ/bmg.c:181-201, 382-410), the step record
// In this kernel sequence no kernel waits on another workgroup: every cross-workgroup step
// is a kernel boundary.  The persistent step (kp_step, opt-in) runs the same items of
// every class in one launch per step instead, ordered by device counters (bit-identical
// results; slower on this chip, see the ticket sizes below).  The random numbers are the reference's own MT19937 stream, taken by index from a
// lookup window built on the host (StzMtWindow), so a resident run draws exactly what
// the per-call path and the reference draw.  A replica halts (every later kernel of
// it exits at once) on termination, on a print step (the host re-solves and writes
// output), when the window runs short, or on error; the host services it between
// batches.  The step sequence is launched directly or replayed from a hipGraph that
// holds a fixed number of consecutive steps.
#include "stz_kmc_dev.h"
#include "stz_local_dev.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ RateArgs<R> rate_args_of(const RepDev &d, const RepState *st, long long N, long long nblk,
                                                    int M)
{
    RateArgs<R> a;
    a.sigma = (R *)d.sigma;
    a.q0 = (const R *)d.q0;
    a.deps = (const R *)d.deps;
    a.deps_last = (const R *)d.deps_last;
    a.q0eff = (R *)d.q0eff;
    a.scache = (R *)d.scache;
    a.perm = d.perm;
    a.temp = d.temp;
    a.tlast = d.tlast;
    a.minq = (R *)d.minq;
    a.mro = nullptr;                 // crystal-like MRO runs take the host path
    a.argm = d.argm;
    a.bsum = d.bsum;
    a.acnt = d.acnt;
    a.nneg = reinterpret_cast<unsigned long long *>(d.res + RES_NNEG);
    a.thr = d.thr;
    a.N = N;
    a.nblk = nblk;
    a.M = M;
    a.softening = d.softening;
    a.pend_comp = st->pend_comp;
    a.pend_delta = st->pend_delta;
    for (int c = 0; c < 6; ++c) a.soff[c] = st->soff[c];
    a.totime = st->totime;
    a.tTemp = d.tTemp;
    a.kT = d.kT;
    a.ratepre = d.ratepre;
    a.voxel = d.voxel;
    a.lnpre = d.lnpre;
    return a;
}

__device__ __forceinline__ ApplyArgs apply_args_of(const RepDev &d, const RepState *st, long long N, int M)
{
    ApplyArgs a;
    a.eps0 = d.eps0;
    a.deps = d.deps;
    a.deps_last = d.deps_last;
    a.scache = d.scache;
    a.perm = d.perm;
    a.temp = d.temp;
    a.tlast = d.tlast;
    a.tloc = d.tloc;
    a.sigma = d.sigma;
    for (int c = 0; c < 6; ++c) a.soff[c] = st->soff[c];
    a.q0 = d.q0;
    a.mro = nullptr;
    a.mro_q0 = 0.0;
    a.rratio = 0.0;
    a.unew = nullptr;
    a.N = N;
    a.M = M;
    a.permsoft = d.permsoft;
    a.tempsoft = d.tempsoft;
    a.permcap = d.permcap;
    a.totime = st->totime;
    return a;
}

// Each kernel of the step is a thin wrapper around an item function (one workgroup's
// share of one replica); the persistent step (kp_step) calls the same item functions.
template <typename R, int DIM>
__device__ __forceinline__ void rates_item(const RepDev &d, long long N, long long nblk, int M, long long blk)
{
    const RepState *st = d.st;
    if (st->halt) return;
    const RateArgs<R> a = rate_args_of<R>(d, st, N, nblk, M);
    rates_block<R, DIM>(a, blk);
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void kb_rates(const RepDev *__restrict__ reps, long long N, long long nblk, int M)
{
    rates_item<R, DIM>(reps[blockIdx.y], N, nblk, M, blockIdx.x);
}

__device__ __forceinline__ void group_item(const RepDev &d, long long nb, long long ngrp, long long gi, double *ws)
{
    if (d.st->halt) return;
    group_sums(d.bsum, nb, d.gsum, ngrp, gi * STZ_GRP_PER_BLOCK, ws);
}

__global__ __launch_bounds__(NT) void kb_group(const RepDev *__restrict__ reps, long long nb, long long ngrp)
{
    __shared__ double ws[NW];
    group_item(reps[blockIdx.y], nb, ngrp, blockIdx.x, ws);
}

// Chunk totals per replica, one workgroup per chunk: the running scan's total of NT
// group sums (chunk c < nchr), and the athermal count of NT voxel blocks (c < ncha).
// Both levels let the single-workgroup decision read a few thousand values at most.
__device__ __forceinline__ void chunks_item(const RepDev &d, long long ngrp, long long nblk, long long c, double *sh,
                                            int *wsi)
{
    if (d.st->halt) return;
    if (c < (ngrp + NT - 1) / NT) {
        const double t = chunk_total(d.gsum, ngrp, c, sh);
        if (threadIdx.x == NT - 1) d.csum[c] = t;
    }
    if (c < (nblk + NT - 1) / NT) {
        const long long b = c * NT + threadIdx.x;
        const int x = wave_sum_int(b < nblk ? d.acnt[b] : 0);
        if (threadIdx.x % WAVE == 0) wsi[threadIdx.x / WAVE] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int i = 0; i < NW; ++i) t += wsi[i];
            d.ccnt[c] = t;
        }
    }
}

__global__ __launch_bounds__(NT) void kb_chunks(const RepDev *__restrict__ reps, long long ngrp, long long nblk)
{
    __shared__ double sh[NW];
    __shared__ int wsi[NW];
    chunks_item(reps[blockIdx.y], ngrp, nblk, blockIdx.x, sh, wsi);
}

// One workgroup per replica: R, the Q<0 count, the step kind, and the stream positions
// of the step's random numbers; then the residence-time selection and transform of a
// thermal event, or the list offsets of an athermal avalanche.  A step whose draws are
// not all inside the lookup window is marked KIND_STALL (nothing changes).
template <typename R, int DIM>
__device__ __forceinline__ void decide_item(const RepDev &d, long long N, long long nblk, long long ngrp, int M)
{
    constexpr int S = Comp<DIM>::S;
    __shared__ int s_kind;
    __shared__ double s_tau;
    __shared__ long long s_k0;
    __shared__ int s_par;
    __shared__ double s_R;
    RepState *st = d.st;
    if (st->halt) return;
    double *res = d.res;
    const int tid = threadIdx.x;
    if (tid == 0) s_R = chunks_total(d.csum, (ngrp + NT - 1) / NT);
    __syncthreads();
    const double Rtot = s_R;
    // rate args for the selection: the pending stress increment is in the field now
    RateArgs<R> a = rate_args_of<R>(d, st, N, nblk, M);
    a.pend_comp = -1;
    a.pend_delta = 0.0;
    const MtDev &mt = d.mt;
    const long long MS = (long long)M * S;
    __syncthreads();
    if (tid == 0) {
        unsigned long long *cnt = reinterpret_cast<unsigned long long *>(res + RES_NNEG);
        const unsigned long long nn = *cnt;
        *cnt = 0ull;
        res[8] = Rtot;
        res[9] = (double)nn;
        st->pend_comp = -1;          // consumed by the rate pass
        st->pend_delta = 0.0;
        reinterpret_cast<int *>(res + 15)[0] = 0;
        res[14] = 0.0;
        res[13] = 0.0;
        const long long rpos = st->rpos;
        int kind;
        if (nn > 0) {
            kind = KIND_ATHERMAL;                        // STZ/bmg.c:161
        } else if (1 / Rtot < d.dt && d.kresp && st->nlocal >= d.local_resync) {
            kind = KIND_NEEDSOLVE;                       // a thermal event, but a full solve is due
        } else if (1 / Rtot < d.dt) {                    // STZ/evolution.c:134-136
            kind = KIND_THERMAL;
            // the dice (STZ/evolution.c:142), then M*S gaussians (STZ/elastic.c:508-512)
            if (rpos - mt.base >= mt.len) {
                kind = KIND_STALL;
            } else {
                const double dice = mt.w[rpos - mt.base] / 4294967296.0;
                int p;
                const long long k0 = mt_first(mt, rpos + 1, p);
                if (k0 + MS > mt.nacc[p]) {
                    kind = KIND_STALL;
                } else if (!(dice > 0.0)) {
                    kind = KIND_FAIL;
                } else {
                    s_tau = dice * Rtot;
                    s_k0 = k0;
                    s_par = p;
                    reinterpret_cast<long long *>(res + RES_RPOS)[0] = mt_after(mt, p, k0 + MS - 1);
                }
            }
        } else {
            kind = KIND_ELASTIC;
        }
        s_kind = kind;
    }
    __syncthreads();
    int kind = s_kind;
    if (kind == KIND_ATHERMAL) {
        // exclusive offsets of the athermal chunks (aoff[c]); kb_athermal derives the
        // block offsets inside each chunk
        const long long n = block_offsets(d.ccnt, (nblk + NT - 1) / NT, d.aoff);
        if (tid == 0) {
            const long long rpos = st->rpos;
            int p = 0;
            long long k0 = mt.nacc[0] + 1;
            if (rpos - mt.base <= mt.len) k0 = mt_first(mt, rpos, p);
            if (d.kresp && (n > d.local_max || st->nlocal >= d.local_resync)) {
                kind = KIND_NEEDSOLVE;                   // the avalanche goes to the full solve
            } else if (k0 + n * MS > mt.nacc[p]) {
                kind = KIND_STALL;
            } else {
                res[13] = (double)n;
                res[RES_K0] = (double)k0;
                reinterpret_cast<int *>(res + RES_PAR)[0] = p;
                reinterpret_cast<long long *>(res + RES_RPOS)[0] = n > 0 ? mt_after(mt, p, k0 + n * MS - 1) : rpos;
            }
            s_kind = kind;
        }
        __syncthreads();
        kind = s_kind;
    } else if (kind == KIND_THERMAL) {
        const ApplyArgs ap = apply_args_of(d, st, N, M);
        const double *src = mt.g[s_par] + s_k0;
        const bool ok = select_block<R, DIM>(a, d.gsum, d.csum, ngrp, s_tau, res, ap, src, d.half_gamma0);
        if (!ok) kind = KIND_FAIL;
    }
    if (tid == 0) {
        // stress after the transforms: the event-local update for a thermal event or a small
        // avalanche while the field is fresh enough, else the full solve
        const bool solved = kind == KIND_THERMAL || kind == KIND_ATHERMAL;
        const int local = (solved && d.kresp) ? 1 : 0;    // the local replica's full solves run on the host
        reinterpret_cast<int *>(res + RES_LOCAL)[0] = local;
        reinterpret_cast<int *>(res + RES_FFT)[0] = (solved && !local) ? kind : KIND_NONE;
        reinterpret_cast<int *>(res + RES_KIND)[0] = kind;
    }
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void kb_decide(const RepDev *__restrict__ reps, long long N, long long nblk,
                                                long long ngrp, int M)
{
    decide_item<R, DIM>(reps[blockIdx.x], N, nblk, ngrp, M);
}

// event-local stress update of the replicas whose step decided it (stz_local_dev.h): the
// chosen site of a thermal event or the avalanche list, each with the strain it recorded
template <typename R, int DIM>
__device__ __forceinline__ void local_item(const RepDev &d, int fp64, int lb)
{
    constexpr int C = Comp<DIM>::C;
    __shared__ LocalSrc<R> src[STZ_LOCAL_MAXSRC];
    __shared__ double wred[NW_LOCAL * 6];
    __shared__ double uni[6];
    const RepState *st = d.st;
    double *res = d.res;
    if (st->halt || reinterpret_cast<const int *>(res + RES_LOCAL)[0] != 1) return;
    const LocalGeom g = local_geom(DIM, d.n, local_row_voxels(sizeof(R) == 4));
    const bool thermal = reinterpret_cast<const int *>(res + RES_KIND)[0] == KIND_THERMAL;
    const int ns = thermal ? 1 : (int)res[13];
    for (int s = threadIdx.x; s < ns; s += NT_LOCAL) {
        const int v = thermal ? (int)res[10] : d.alist[s];
        load_src<R, DIM>(g, (const R *)d.deps_last, v, src[s]);
    }
    if (threadIdx.x == 0) {
        double de[6], u[6];
        for (int c = 0; c < 6; ++c) {
            de[c] = st->epsavg[c] - st->ebase[c];
            u[c] = 0.0;
        }
        c0_mul<C>(de, d.lambda, d.mu, u);
        for (int c = 0; c < 6; ++c) uni[c] = fp64 ? u[c] - st->usum[c] : u[c];
    }
    __syncthreads();
    local_rows<R, DIM>(g, (const R *)d.kresp, src, ns, uni, (R *)d.sigma, lb, d.spart,
                       reinterpret_cast<int *>(res + 15), wred);
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT_LOCAL) void kb_local(const RepDev *__restrict__ reps, int fp64)
{
    local_item<R, DIM>(reps[blockIdx.y], fp64, blockIdx.x);
}

// Athermal compaction and transforms: one workgroup per chunk of NT voxel blocks.  The
// rate pass counted the flagged voxels of every block (acnt) and kb_chunks those of every
// chunk (ccnt); kb_decide turned the chunk counts into offsets (aoff[c]).  A workgroup
// scans its chunk's block counts for the block offsets and works through the blocks that
// have flagged voxels -- usually none, so most workgroups read NT counts and return, and
// the min-Q field is read only where an avalanche is.  The partials socp[blk] are written
// for those blocks only; kb_finish sums the others as the +0.0 they are.
template <typename R, int DIM>
__device__ __forceinline__ void athermal_item(const RepDev &d, long long N, int M, long long nblk, long long c)
{
    __shared__ int sw[NW];
    __shared__ int s_nb;
    __shared__ long long s_blk[NT], s_off[NT];
    const RepState *st = d.st;
    const double *res = d.res;
    if (st->halt || reinterpret_cast<const int *>(res + RES_KIND)[0] != KIND_ATHERMAL) return;
    if (d.ccnt[c] == 0) return;                 // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    const long long blk = c * NT + tid;
    const int cnt = blk < nblk ? d.acnt[blk] : 0;
    // exclusive scan of the block counts (exact integers) and the list of non-empty blocks
    int x = cnt;
#pragma unroll
    for (int dd = 1; dd < WAVE; dd <<= 1) {
        const int y = __shfl_up(x, dd, WAVE);
        if (lane >= dd) x += y;
    }
    if (lane == WAVE - 1) sw[w] = x;
    const unsigned long long ne = __ballot(cnt != 0);
    __syncthreads();
    int woff = 0;
    for (int i = 0; i < w; ++i) woff += sw[i];
    const long long boff = (long long)d.aoff[c] + woff + (x - cnt);
    __shared__ int sne[NW];
    if (lane == 0) sne[w] = __popcll(ne);
    __syncthreads();
    int nbefore = 0, ntot = 0;
    for (int i = 0; i < NW; ++i) {
        if (i < w) nbefore += sne[i];
        ntot += sne[i];
    }
    if (cnt != 0) {
        const int slot = nbefore + __popcll(ne & ((1ull << lane) - 1ull));
        s_blk[slot] = blk;
        s_off[slot] = boff;
    }
    if (tid == 0) s_nb = ntot;
    __syncthreads();
    AthArgs<R> a;
    a.ap = apply_args_of(d, st, N, M);
    a.minq = (const R *)d.minq;
    a.argm = d.argm;
    a.sigma = (const R *)d.sigma;
    for (int cc = 0; cc < 6; ++cc) a.soff[cc] = st->soff[cc];
    a.voxel = d.voxel;
    a.thr = d.thr;
    a.list = d.alist;
    a.off = d.aoff;
    const int p = reinterpret_cast<const int *>(res + RES_PAR)[0];
    a.dsrc = d.mt.g[p] + (long long)res[RES_K0];
    a.dscale = d.half_gamma0;
    a.cap = 1ll << 62;
    a.N = N;
    a.socp = d.socp;
    const int nb = s_nb;
    int par = 0;
    for (int i = 0; i < nb; ++i) {
        const long long b = s_blk[i];
        const long long v = b * NT + tid;
        athermal_block_f<R, DIM>(a, b, v < N ? athermal_flag(a, v) : 0, par, s_off[i]);
        par ^= 1;
    }
}

template <typename R, int DIM>
__global__ __launch_bounds__(NT) void kb_athermal(const RepDev *__restrict__ reps, long long N, int M,
                                                  long long nblk)
{
    athermal_item<R, DIM>(reps[blockIdx.y], N, M, nblk, blockIdx.x);
}

// --- host-loop arithmetic of the reference, on the device, with the host's operation order
#pragma clang fp contract(off)

// a[i] and a[i] += v for a runtime component i (the loaded component) as unrolled
// selects: a private copy of the run state indexed at run time is placed in scratch
// memory, and every access of the bookkeeping then is a memory round trip
__device__ __forceinline__ double get6(const double *a, int i)
{
    double r = a[0];
#pragma unroll
    for (int c = 1; c < 6; ++c) {
        double x = a[c];
        asm volatile("" : "+v"(x));   // an opaque value: the select is not folded back into
        r = (i == c) ? x : r;         // one load through a runtime address
    }
    return r;
}
__device__ __forceinline__ void add6(double *a, int i, double v)
{
#pragma unroll
    for (int c = 0; c < 6; ++c)
        if (c == i) a[c] += v;
}

// isotropicinv (STZ/elastic.c:522-531)
template <int C>
__device__ __forceinline__ void iso_inv(const RepDev &d, const double *st, double *e)
{
    const double lam = d.lambda, mu = d.mu;
    if constexpr (C == 3) {
        double tra = -lam / 2 / mu / (2 * lam + 2 * mu) * (st[0] + st[1]);
        e[0] = tra + st[0] / 2 / mu;
        e[1] = tra + st[1] / 2 / mu;
        e[2] = st[2] / 2 / mu;
    } else {
        double tra = -lam / 2 / mu / (3 * lam + 2 * mu) * ((st[0] + st[1]) + st[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i) e[i] = tra + st[i] / 2 / mu;
#pragma unroll
        for (int i = 3; i < 6; ++i) e[i] = st[i] / 2 / mu;
    }
}

// Sigma[comp][*] += delta: FP64 defers it to the next rate pass, mixed keeps an offset
__device__ __forceinline__ void add_uniform_d(RepState *s, bool fp64, int comp, double delta)
{
    if (fp64) {
        s->pend_comp = comp;
        s->pend_delta = delta;
        add6(s->usum, comp, delta);  // in the field once the next rate pass applies it
    } else {
        add6(s->soff, comp, delta);
    }
}

// KMC loading after a thermal event (STZ/evolution.c:281-315)
template <int C>
__device__ __forceinline__ void load_thermal_d(const RepDev &d, RepState *s, bool fp64, double t)
{
    const double sr = d.strainrate, mu = d.mu;
    const int lc = d.lc;
    double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
    if (d.loading == STZ_SHEAR) {
#pragma unroll
        for (int c = 0; c < C; ++c) ts[c] = (c == lc) ? 0 : s->stress[c];
        iso_inv<C>(d, ts, te);
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (c != lc) s->epsavg[c] += -te[c];
        add6(s->epsavg, lc, sr * t);
        add6(s->stress, lc, 2 * mu * sr * t);
        add_uniform_d(s, fp64, lc, 2 * mu * sr * t);
        return;
    }
    const double nu = d.nu, E = d.E;
#pragma unroll
    for (int c = 1; c < C; ++c) ts[c] = s->stress[c];
    iso_inv<C>(d, ts, te);
    s->epsavg[0] += (sr * t * 1 - te[0]);
    s->epsavg[1] += (sr * t * (-nu) - te[1]);
    if constexpr (C == 3) {
        s->epsavg[2] += (-te[2]);
    } else {
        s->epsavg[2] += (sr * t * (-nu) - te[2]);
#pragma unroll
        for (int c = 3; c < 6; ++c) s->epsavg[c] += (-te[c]);
    }
    s->stress[0] += E * sr * t;
    add_uniform_d(s, fp64, 0, E * sr * t);
}

// pure elastic increment (STZ/evolution.c:321-358)
template <int C>
__device__ __forceinline__ void load_elastic_d(const RepDev &d, RepState *s, bool fp64)
{
    const double sr = d.strainrate, inc = d.strainincrement, mu = d.mu;
    const int lc = d.lc;
    if (d.loading == STZ_SHEAR) {
        add6(s->epsavg, lc, sr * 1);
        add6(s->stress, lc, 2 * mu * inc);
        add_uniform_d(s, fp64, lc, 2 * mu * inc);
        return;
    }
    const double nu = d.nu, E = d.E;
    s->epsavg[0] += (inc * 1);
    s->epsavg[1] += (inc * (-nu));
    if constexpr (C == 3) {
        s->epsavg[2] += (inc * 0.0);
    } else {
        s->epsavg[2] += (inc * (-nu));
#pragma unroll
        for (int c = 3; c < 6; ++c) s->epsavg[c] += (inc * 0.0);
    }
    s->stress[0] += E * inc;
    add_uniform_d(s, fp64, 0, E * inc);
}

// athermal error subtraction, STZ/evolution.c:91-97 (reference) or loading-consistent
template <int C>
__device__ __forceinline__ void athermal_correction_d(const RepDev &d, RepState *s)
{
    double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
    if (d.athermal_correction == 0) {
#pragma unroll
        for (int c = 1; c < C; ++c) ts[c] = s->stress[c];
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c) ts[c] = (c == d.lc) ? 0.0 : s->stress[c];
    }
    iso_inv<C>(d, ts, te);
#pragma unroll
    for (int c = 0; c < C; ++c) s->epsavg[c] -= te[c];
}

__device__ __forceinline__ bool terminated_d(const RepDev &d, const RepState *s)
{
    if (fabs(get6(s->epsavg, d.lc)) > s->endstrain) return true;                    // bmg.c:383,397
    if (d.loading == STZ_TENSION && s->stress[0] < -0.5) return true;               // bmg.c:388
    if (d.loading == STZ_SHEAR && fabs(get6(s->stress, d.lc)) < STZ_EPSILON) return true; // bmg.c:402
    return false;
}

// tells the host (polling mapped memory while it launches) that the replica halted
__device__ __forceinline__ void publish_halt(const RepDev &d, int h)
{
    __hip_atomic_store(d.hflag, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// the host loop's bookkeeping for one replica, after the device decided and ran its step:
// <sigma> from the solve partials and the |DWork| sum (all threads, the per-call path's
// reductions), then the scalar updates on a private copy of the run state (one thread)
// NTH: the workgroup size; the reductions are the STZ_NT_MEAN-thread trees whatever it is
template <int DIM, int NTH>
__device__ __forceinline__ void finish_item(const RepDev &d, int fp64, int nparts, int nparts_local, double invN,
                                            long long nblk)
{
    constexpr int C = Comp<DIM>::C;
    __shared__ double mean[C];
    RepState *sp = d.st;
    if (sp->halt) return;
    double *res = d.res;
    int *kindp = reinterpret_cast<int *>(res + RES_KIND);
    const int kind = kindp[0];
    const bool solved = kind == KIND_THERMAL || kind == KIND_ATHERMAL;
    const int local = reinterpret_cast<const int *>(res + RES_LOCAL)[0];
    double soc = 0.0;
    if constexpr (NTH == STZ_NT_MEAN) {
        if (solved) mean_partials<C>(d.spart, local ? nparts_local : nparts, invN, mean);
        if (kind == KIND_ATHERMAL) soc = ordered_sum_flagged<STZ_NT_MEAN>(d.socp, d.acnt, d.ccnt, nblk);
    } else {
        if (solved) mean_partials_n<C, NTH>(d.spart, local ? nparts_local : nparts, invN, mean);
        if (kind == KIND_ATHERMAL) soc = ordered_sum_flagged_n<STZ_NT_MEAN, NTH>(d.socp, d.acnt, d.ccnt, nblk);
    }
    if (threadIdx.x != 0) return;
    // the solve gates of this step are spent (every thread has read them: the reductions
    // above end in barriers, and without a reduction no other thread uses them)
    reinterpret_cast<int *>(res + RES_FFT)[0] = KIND_NONE;
    reinterpret_cast<int *>(res + RES_LOCAL)[0] = 0;
    RepState s = *sp;
    if (kind == KIND_STALL || kind == KIND_NEEDSOLVE) {
        const int h = kind == KIND_STALL ? HALT_RNG : HALT_SOLVE;
        sp->halt = h;
        publish_halt(d, h);
        kindp[0] = KIND_NONE;
        return;
    }
    if (kind == KIND_FAIL || (kind != KIND_THERMAL && kind != KIND_ATHERMAL && kind != KIND_ELASTIC)) {
        sp->halt = HALT_ERR;
        sp->err = kind == KIND_FAIL ? STZ_E_SELECT : STZ_E_HIP;
        publish_halt(d, HALT_ERR);
        kindp[0] = KIND_NONE;
        return;
    }
    if (solved && reinterpret_cast<const int *>(res + 15)[0]) {
        sp->halt = HALT_ERR;
        sp->err = STZ_E_NAN;
        publish_halt(d, HALT_ERR);
        kindp[0] = KIND_NONE;
        return;
    }
    const double R = res[8];
    stz_step_result r;
    r.kind = STZ_STEP_ELASTIC;
    r.site = r.mode = -1;
    r.n_transformed = 0;
    r.n_negative = (long long)res[9];
    r.q_chosen = 0.0;
    r.total_rate = R;
    r.dt_kmc = 0.0;
    r.printed = 0;
    r.soc_elast = 0.0;
    s.step++;
    if (solved) {
        s.nsolve++;
        for (int c = 0; c < C; ++c) s.stress[c] = mean[c];
        for (int c = 0; c < 6; ++c) {
            s.soff[c] = 0.0;                  // the solve (or update) rewrote the stress field
            s.ebase[c] = s.epsavg[c];         // with the mean strain of this step
            s.usum[c] = 0.0;
        }
        s.nlocal = local ? s.nlocal + 1 : 0;
        s.rpos = reinterpret_cast<const long long *>(res + RES_RPOS)[0];
    }
    if (kind == KIND_THERMAL) {
        s.avlan = 0;                          // UpdateSOC (STZ/evolution.c:5-45)
        s.soc_elast = 0.0;
        const double t = 1 / R;               // STZ/evolution.c:134
        r.dt_kmc = t;
        s.kstep++;
        s.ntrans++;
        r.kind = STZ_STEP_THERMAL;
        r.site = (int)res[10];
        r.mode = (int)res[11];
        r.q_chosen = res[12];
        load_thermal_d<C>(d, &s, fp64 != 0, t);
        s.totime += t;                        // STZ/evolution.c:317
    } else if (kind == KIND_ATHERMAL) {
        s.astep++;                            // athermal_plasticity (STZ/evolution.c:47-100)
        const long long n = (long long)res[13];
        s.ntrans += n;
        s.avlan += n;
        s.soc_elast += soc;
        athermal_correction_d<C>(d, &s);
        r.kind = STZ_STEP_ATHERMAL;
        r.n_transformed = (int)n;
        r.soc_elast = soc;
    } else {
        s.avlan = 0;
        s.soc_elast = 0.0;
        r.dt_kmc = 1 / R;
        s.estep++;
        s.totime += d.dt;                     // STZ/evolution.c:324
        load_elastic_d<C>(d, &s, fp64 != 0);
    }
    if (s.print_on) {
        // STZ/bmg.c:181-187; a period <= 0 makes the test true at every step, as there
        const double cur = get6(s.epsavg, d.lc);
        if (fabs(cur - s.strainhist_last) >= (s.printperiod * d.strainincrement * 0.99)) {
            s.strainhist_last = cur;          // strainhist[N_print-1], STZ/bmg.c:193-201
            r.printed = 1;
            s.halt = HALT_PRINT;
        } else if (terminated_d(d, &s)) {
            s.halt = HALT_DONE;
        }
    }
    r.totime = s.totime;
    for (int c = 0; c < 6; ++c) {
        r.epsavg[c] = s.epsavg[c];
        r.stress[c] = s.stress[c];
    }
    d.rec[s.nrec] = r;
    s.nrec++;
    if (s.halt == HALT_NONE && s.nrec >= s.budget) s.halt = HALT_BUDGET;
    *sp = s;
    if (s.halt != HALT_NONE) publish_halt(d, s.halt);
    kindp[0] = KIND_NONE;
}

template <int DIM>
__global__ __launch_bounds__(STZ_NT_MEAN) void kb_finish(const RepDev *__restrict__ reps, int fp64, int nparts,
                                                         int nparts_local, double invN, long long nblk)
{
    finish_item<DIM, STZ_NT_MEAN>(reps[blockIdx.x], fp64, nparts, nparts_local, invN, nblk);
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------------
// Persistent step: one launch runs the whole step of every replica of a plan (all shape
// classes).  The step's work is a list of tickets in dependency order -- phase by phase
// (rates, group sums, chunk totals, decision, athermal pass, local update, bookkeeping),
// within a phase the classes largest first, replica by replica, a few items per ticket.
// The workgroup of a ticket waits until its replica has completed every item of the
// earlier phases (completion counts per replica), runs the items with the item functions
// of the kernels above, and adds them to the count.  Every item a ticket waits for carries
// a smaller ticket: the step progresses without any co-residency promise; the wait is
// bounded in time all the same (abort flag, then every workgroup finishes at once).  The
// replica's last item (its bookkeeping) clears its counts for the next launch.
// Results equal the kernel sequence bit for bit: same item code, same reductions
// (kb_finish's 1024-thread trees computed by 256 threads).
// ---------------------------------------------------------------------------------
constexpr int PH_RATES = 0, PH_GROUP = 1, PH_CHUNKS = 2, PH_DECIDE = 3, PH_ATH = 4, PH_LOCAL = 5, PH_FINISH = 6;
constexpr int NPH = 7;
constexpr int PMAXCLS = 8;
// sync words: the abort flag, then per replica STZ_PERSIST_SLOTS completion counts, each
// on a 128-byte line of its own (a ticket adds its items to slot ticket % slots, a waiter
// sums the slots), so the completions of one phase do not all update one address
// (measured: the same step time with one slot; the write-backs dominate)
#ifndef STZ_PERSIST_SLOTS
#define STZ_PERSIST_SLOTS 16
#endif
constexpr int PSYNC_ABORT = 0, PSYNC_REP = 32, PSYNC_LINE = 32;
constexpr int PSYNC_PER_REP = STZ_PERSIST_SLOTS * PSYNC_LINE;
__host__ __device__ constexpr long long psync_words(int nrep) { return PSYNC_REP + (long long)nrep * PSYNC_PER_REP; }

struct PClass {
    long long N, nblk, ngrp;
    long long items[NPH];         // items of each phase, per replica
    long long before[NPH];        // items of the earlier phases, per replica
    int per[NPH];                 // items per ticket
    int M, rep0, nparts_local;    // rep0: the first replica's completion count
    double invN;
};
struct PSeg {
    unsigned t0, t1;              // tickets [t0, t1) of one (phase, class)
    unsigned units;               // tickets per replica
    int ph, cls;
};
// The launch block in device memory: this header, then the RepDev of every replica (class
// by class).  All of it is read through the kernel's one __restrict__ argument, so the
// compiler takes the descriptors with scalar loads into scalar registers, as the kernels
// above do with their reps argument.
struct PArgs {
    PClass cls[PMAXCLS];
    PSeg seg[NPH * PMAXCLS];
    int nseg, fp64, nrep;
    unsigned total;               // tickets of one step
    unsigned *sync;               // [psync_words(nrep)] words, zero between launches
    unsigned long long spin_ticks;   // wall-clock ticks a dependency wait may last
};

// Polling of a waiting workgroup backs off to STZ_PERSIST_BACKOFF sleeps of 512 cycles
// between reads (measured: no change from a one-sleep poll; kept, it costs nothing).
#ifndef STZ_PERSIST_BACKOFF
#define STZ_PERSIST_BACKOFF 16
#endif
// Items per ticket.  Every ticket (workgroup) ends in an agent-scope release, i.e. a
// write-back of its XCD's L2 (8 L2s that are not coherent with each other), where the
// kernel sequence pays one per kernel boundary: fewer, larger tickets cost less, up to
// where the last tickets of a phase leave compute units idle.  Measured per step at
// 512x512x256 mixed (kernel sequence 7.9 ms): 4 rate blocks and 2 row groups per ticket
// 26.4 ms, 16 and 8 13.0 ms, 32 and 16 11.1 ms; the 64-replica ensemble of 32^3, 64x64x128,
// 96^3 and 128^3 grids (7.35 ms): 47.2, 26.7, 21.4 ms.
#ifndef STZ_PERSIST_RPER
#define STZ_PERSIST_RPER 32        // rate blocks per ticket
#endif
#ifndef STZ_PERSIST_LPER
#define STZ_PERSIST_LPER 16        // local-update row groups per ticket
#endif

// ticket t (< total) -> phase, class, replica within the class, items [i0, i1) of the phase
template <typename PA>
__host__ __device__ __forceinline__ void persist_decode(PA *P, unsigned t, int &ph, int &ci, int &r, long long &i0,
                                                        long long &i1)
{
    int si = 0;
    while (t >= P->seg[si].t1) ++si;
    ph = P->seg[si].ph;
    ci = P->seg[si].cls;
    const unsigned units = P->seg[si].units;
    const unsigned u = t - P->seg[si].t0;
    r = (int)(u / units);
    const long long per = P->cls[ci].per[ph], n = P->cls[ci].items[ph];
    i0 = (long long)(u % units) * per;
    i1 = i0 + per < n ? i0 + per : n;
}

// the shape of one class as the table needs it
struct PShape {
    long long N, nblk, ngrp, nlb;
    int M, nrep;
};

// Builds the ticket table of classes shp[0..n) (already in launch order, largest first):
// items per phase and replica, the segments of (phase, class), the replicas' completion
// counts.  false when it does not fit (more than PMAXCLS classes, counts beyond 31 bits).
inline bool persist_table(const PShape *shp, int n, PArgs &A)
{
    static const int per[NPH] = {STZ_PERSIST_RPER, 1, 1, 1, 8, STZ_PERSIST_LPER, 1};
    if (n < 1 || n > PMAXCLS) return false;
    std::memset(&A, 0, sizeof(A));
    int nrep = 0;
    for (int j = 0; j < n; ++j) {
        const PShape &p = shp[j];
        PClass &c = A.cls[j];
        c.N = p.N;
        c.nblk = p.nblk;
        c.ngrp = p.ngrp;
        c.M = p.M;
        c.rep0 = nrep;
        c.nparts_local = (int)p.nlb;
        c.invN = 1.0 / (double)p.N;
        c.items[PH_RATES] = p.nblk;
        c.items[PH_GROUP] = (p.ngrp + STZ_GRP_PER_BLOCK - 1) / STZ_GRP_PER_BLOCK;
        c.items[PH_CHUNKS] = std::max((p.ngrp + NT - 1) / NT, (p.nblk + NT - 1) / NT);
        c.items[PH_DECIDE] = 1;
        c.items[PH_ATH] = (p.nblk + NT - 1) / NT;
        c.items[PH_LOCAL] = p.nlb;
        c.items[PH_FINISH] = 1;
        long long b = 0;
        for (int ph = 0; ph < NPH; ++ph) {
            c.per[ph] = per[ph];
            c.before[ph] = b;
            b += c.items[ph];
        }
        if (b >= (1ll << 31)) return false;      // beyond the 32-bit completion counts
        nrep += p.nrep;
    }
    long long t = 0;
    int ns = 0;
    for (int ph = 0; ph < NPH; ++ph)
        for (int j = 0; j < n; ++j) {
            const PClass &c = A.cls[j];
            const long long units = (c.items[ph] + c.per[ph] - 1) / c.per[ph];
            if (units == 0) continue;
            PSeg &g = A.seg[ns++];
            g.t0 = (unsigned)t;
            t += units * (long long)shp[j].nrep;
            g.t1 = (unsigned)t;
            g.units = (unsigned)units;
            g.ph = ph;
            g.cls = j;
        }
    if (t >= (1ll << 31)) return false;          // one workgroup per ticket: a grid dimension
    A.nseg = ns;
    A.total = (unsigned)t;
    A.nrep = nrep;
    return true;
}

// the replica's completed items: the sum of its slots
__device__ __forceinline__ unsigned count_sum(const unsigned *cnt)
{
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < STZ_PERSIST_SLOTS; ++i)
        s += __hip_atomic_load(cnt + i * PSYNC_LINE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return s;
}

// thread 0: waits until the replica's completed items reach need, then acquires (the
// writes of those items are visible to the workgroup); false when the step aborts
__device__ __forceinline__ bool dep_wait(const unsigned *cnt, unsigned need, unsigned *abortf, unsigned long long limit)
{
    if (count_sum(cnt) < need) {
        const unsigned long long t0 = wall_clock64();
        unsigned nap = 1;
        for (;;) {
            for (unsigned j = 0; j < nap; ++j) __builtin_amdgcn_s_sleep(8);
            if (nap < STZ_PERSIST_BACKOFF) nap *= 2;
            if (count_sum(cnt) >= need) break;
            if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
            if (wall_clock64() - t0 > limit) {
                __hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// The launch block is read-only while a launch runs, so it is read through the constant
// address space: uniform loads of it (and of the descriptors) are scalar loads into
// scalar registers, which the compiler cannot otherwise prove safe next to the phases'
// stores.
typedef __attribute__((address_space(4))) const PArgs CPArgs;
typedef __attribute__((address_space(4))) const RepDev CRepDev;

// One workgroup per ticket, ticket = block index.  Workgroups are dispatched in block-index
// order, so every ticket a workgroup waits for belongs to a workgroup dispatched before it,
// which runs to its end: the waits cannot close a cycle, whatever the number of resident
// workgroups.  (Tickets taken from an atomic counter instead, with an exit counter for
// the reset, cost about 260 ns per ticket in same-address atomics: 31 ms per step at
// 512x512x256 against 8 ms for the kernel sequence.  A loop over tickets in fewer
// workgroups let the compiler hoist every phase's loop-invariant values out of the loop
// and keep them all live: 250 vector registers, one wave per SIMD.)  Three waves per SIMD
// at least: the FP64 3-D variant then fits in 165 registers without spilling, like the
// FP64 rate pass and local update alone; the others reach four.
template <typename R, int DIM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void kp_step(const PArgs *__restrict__ Pg)
{
    __shared__ int s_ok;
    __shared__ double ws[NW];
    __shared__ int wsi[NW];
    const CPArgs *P = (const CPArgs *)(unsigned long long)Pg;
    unsigned *sync = P->sync;
    const unsigned t = blockIdx.x;
    int ph, ci, r;
    long long i0, i1;
    persist_decode(P, t, ph, ci, r, i0, i1);
    const auto &k = P->cls[ci];
    unsigned *cnt = sync + PSYNC_REP + (long long)(k.rep0 + r) * PSYNC_PER_REP;
    if (threadIdx.x == 0) {
        // after an abort no workgroup runs its items any more (they would read the
        // unfinished results of the items that skipped theirs)
        const bool ok = k.before[ph] == 0 || dep_wait(cnt, (unsigned)k.before[ph], sync + PSYNC_ABORT, P->spin_ticks);
        s_ok = ok && __hip_atomic_load(sync + PSYNC_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
    }
    // thread 0's acquire invalidated the compute unit's vector caches, which every wave of
    // the workgroup reads through after the barrier
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s_ok)) {
        const RepDev &d = *(const RepDev *)((const CRepDev *)(unsigned long long)(Pg + 1) + k.rep0 + r);
        const long long N = k.N, nblk = k.nblk, ngrp = k.ngrp;
        const int M = k.M, fp64 = P->fp64;
        auto items = [&](auto f) {
            for (long long i = i0; i < i1; ++i) {
                f(i);
                __syncthreads();
            }
        };
        switch (ph) {
        case PH_RATES: items([&](long long i) { rates_item<R, DIM>(d, N, nblk, M, i); }); break;
        case PH_GROUP: items([&](long long i) { group_item(d, (long long)M * nblk, ngrp, i, ws); }); break;
        case PH_CHUNKS: items([&](long long i) { chunks_item(d, ngrp, nblk, i, ws, wsi); }); break;
        case PH_DECIDE: decide_item<R, DIM>(d, N, nblk, ngrp, M); break;
        case PH_ATH: items([&](long long i) { athermal_item<R, DIM>(d, N, M, nblk, i); }); break;
        case PH_LOCAL: items([&](long long i) { local_item<R, DIM>(d, fp64, (int)i); }); break;
        default: finish_item<DIM, NT>(d, fp64, 0, k.nparts_local, k.invN, nblk); break;
        }
    }
    // every wave's writes reach agent scope before the count moves
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ph == PH_FINISH) {
            // the replica's last item: every other item of it has completed and no workgroup
            // of this launch reads its counts any more; they start from zero next launch
            for (int i = 0; i < STZ_PERSIST_SLOTS; ++i)
                __hip_atomic_store(cnt + i * PSYNC_LINE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_fetch_add(cnt + (t % STZ_PERSIST_SLOTS) * PSYNC_LINE, (unsigned)(i1 - i0), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace

// ==================================================================================
// host side
// ==================================================================================
struct StzResident {
    RepState *st = nullptr;           // device
    stz_step_result *rec = nullptr;   // device [rec_cap]
    int rec_cap = 0;
    // MT19937 lookup window (device copies of StzMtWindow's arrays)
    uint32_t *w = nullptr;
    double *g[2] = {nullptr, nullptr};
    int *rank[2] = {nullptr, nullptr};
    int *pair[2] = {nullptr, nullptr};
    long long wcap = 0;               // words the device arrays hold
    long long wsize = 1ll << 21;      // words per window
    StzMtWindow win;
    bool win_ok = false;
    long long win_epoch = -1;
    bool force_rebuild = false;
    bool rng_lazy = false;            // ctx->rng lags behind the resident stream position
    long long rng_target = 0;
    RepDev hdev;                      // host copy of the replica descriptor
    RepState hst;                     // host copy of the run state
    std::vector<stz_step_result> hrec;
    int batch = 64;
    int stop_at_print = 0;
    int threads = 8;
    int *hflag = nullptr;             // mapped pinned halt flag (host address)
    int *hflag_dev = nullptr;         // its device address
    struct StzPlan *self = nullptr;   // single-context plan
};

namespace {

struct StzClass {
    std::vector<stz_ctx *> reps;
    bool local = false;               // some replica updates the stress event-locally
    bool fft = true;                  // some replica takes the full solve on the device
    RepDev *d_reps = nullptr;
    SolveRep *d_solve = nullptr;
    // plans of several classes: the class's own stream (default priority) and the event
    // that joins it back into the plan's stream after each step
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

}  // namespace

struct StzPlan {
    std::vector<StzClass> cls;
    std::vector<stz_ctx *> all;
    hipStream_t stream = nullptr;
    stz_ctx *owner = nullptr;         // context whose error string receives messages
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;        // several classes: their steps run concurrently
    int batch = 64;
    int use_graph = 0;                // steps per captured graph; 0: direct launches
    int graph_steps = 0;              // steps in the instantiated graph
    long long graph_epoch = -1;       // tile choices the graph was captured with
    // rate-pass timing of direct launches (single context with timing on)
    std::vector<hipEvent_t> eva, evb;
    std::vector<hipEvent_t> evc;      // progress markers of the chunked launches
    int timed = 0;
    // persistent step (kp_step): requested, and in use for the current replicas
    int persist = 0;
    bool pactive = false;
    unsigned *psync = nullptr;        // device sync words
    long long psync_n = 0;
    PArgs *pargs = nullptr;           // device launch block: PArgs + the replicas' RepDev
    int pargs_rep = 0;                // replicas it has room for
    PArgs hargs;                      // host copy of the header
    std::vector<RepDev> hreps;        // host copy of the descriptors, class by class
};

namespace {

#define RS_HIP(ctx, call)                                                          \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return STZ_E_HIP;                                                      \
        }                                                                          \
    } while (0)

int rs_alloc(stz_ctx *ctx, void **p, size_t bytes)
{
    RS_HIP(ctx, hipMalloc(p, bytes));
    RS_HIP(ctx, hipMemset(*p, 0, bytes));
    return STZ_OK;
}

void free_window(StzResident *rs)
{
    void *ptrs[] = {rs->w, rs->g[0], rs->g[1], rs->rank[0], rs->rank[1], rs->pair[0], rs->pair[1]};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    rs->w = nullptr;
    for (int p = 0; p < 2; ++p) rs->g[p] = nullptr, rs->rank[p] = nullptr, rs->pair[p] = nullptr;
    rs->wcap = 0;
}

int ensure_resident(stz_ctx *ctx)
{
    if (ctx->rs) return STZ_OK;
    StzResident *rs = new StzResident();
    ctx->rs = rs;
    unsigned hw = std::thread::hardware_concurrency();
    rs->threads = (int)std::max(1u, std::min(hw, 16u));
    int rc = rs_alloc(ctx, (void **)&rs->st, sizeof(RepState));
    if (rc) return rc;
    RS_HIP(ctx, hipHostMalloc((void **)&rs->hflag, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
    RS_HIP(ctx, hipHostGetDevicePointer((void **)&rs->hflag_dev, rs->hflag, 0));
    *rs->hflag = 0;
    rs->rec_cap = rs->batch;
    rc = rs_alloc(ctx, (void **)&rs->rec, sizeof(stz_step_result) * (size_t)rs->rec_cap);
    if (rc) return rc;
    rs->hrec.resize((size_t)rs->rec_cap);
    return STZ_OK;
}

// builds the lookup window that starts at the current position of `from` and uploads it
int window_build(stz_ctx *ctx, const StzRng &from)
{
    StzResident *rs = ctx->rs;
    rs->win.build(from, rs->wsize, rs->threads);
    if (rs->wcap < rs->wsize) {
        free_window(rs);
        const size_t T = (size_t)rs->wsize, H = T / 2 + 1;
        int rc = STZ_OK;
        rc = rc ? rc : rs_alloc(ctx, (void **)&rs->w, T * sizeof(uint32_t));
        for (int p = 0; p < 2; ++p) {
            rc = rc ? rc : rs_alloc(ctx, (void **)&rs->g[p], H * sizeof(double));
            rc = rc ? rc : rs_alloc(ctx, (void **)&rs->rank[p], (H + 1) * sizeof(int));
            rc = rc ? rc : rs_alloc(ctx, (void **)&rs->pair[p], H * sizeof(int));
        }
        if (rc) return rc;
        rs->wcap = rs->wsize;
    }
    const StzMtWindow &W = rs->win;
    RS_HIP(ctx, hipMemcpy(rs->w, W.words.data(), (size_t)W.len * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (int p = 0; p < 2; ++p) {
        if (W.nacc[p] > 0) {
            RS_HIP(ctx, hipMemcpy(rs->g[p], W.g[p].data(), (size_t)W.nacc[p] * sizeof(double), hipMemcpyHostToDevice));
            RS_HIP(ctx, hipMemcpy(rs->pair[p], W.pair[p].data(), (size_t)W.nacc[p] * sizeof(int), hipMemcpyHostToDevice));
        }
        RS_HIP(ctx, hipMemcpy(rs->rank[p], W.rank[p].data(), W.rank[p].size() * sizeof(int), hipMemcpyHostToDevice));
    }
    MtDev &m = rs->hdev.mt;
    m.w = rs->w;
    for (int p = 0; p < 2; ++p) {
        m.g[p] = rs->g[p];
        m.rank[p] = rs->rank[p];
        m.pair[p] = rs->pair[p];
        m.nacc[p] = W.nacc[p];
    }
    m.base = W.base;
    m.len = W.len;
    rs->win_ok = true;
    rs->win_epoch = ctx->rng_epoch;
    return STZ_OK;
}

void fill_desc(stz_ctx *ctx)
{
    StzResident *rs = ctx->rs;
    RepDev &d = rs->hdev;
    d.eps0 = ctx->eps0;
    d.sigma = ctx->sigma;
    d.q0 = ctx->q0;
    d.deps = ctx->deps;
    d.deps_last = ctx->deps_last;
    d.q0eff = ctx->q0eff;
    d.scache = ctx->scache;
    d.minq = ctx->minq;
    d.spec = ctx->spec;
    d.perm = ctx->perm;
    d.temp = ctx->temp;
    d.tlast = ctx->tlast;
    d.tloc = ctx->tloc;
    d.argm = ctx->argm;
    d.bsum = ctx->bsum;
    d.gsum = ctx->gsum;
    d.csum = ctx->csum;
    d.ccnt = ctx->ccnt;
    d.socp = ctx->socp;
    d.spart = ctx->spart;
    d.res = ctx->res;
    d.acnt = ctx->acnt;
    d.aoff = ctx->aoff;
    d.alist = ctx->alist;
    d.st = rs->st;
    d.rec = rs->rec;
    d.hflag = rs->hflag_dev;
    const StzConsts &k = ctx->k;
    d.kT = k.kT;
    d.tTemp = k.tTemp;
    d.ratepre = k.ratepre;
    d.lnpre = std::log(k.ratepre);
    d.voxel = k.voxel;
    d.thr = k.thr;
    d.dt = k.dt;
    d.permsoft = k.permsoft;
    d.tempsoft = k.tempsoft;
    d.permcap = k.permcap;
    d.half_gamma0 = k.half_gamma0;
    d.E = k.E;
    d.nu = k.nu;
    d.mu = k.mu;
    d.lambda = k.lambda;
    d.strainrate = ctx->p.strainrate;
    d.strainincrement = ctx->p.strainincrement;
    d.loading = ctx->p.loading;
    d.lc = ctx->lc;
    d.athermal_correction = ctx->p.athermal_correction;
    d.softening = ctx->p.softening == 1;
    for (int a = 0; a < 3; ++a) d.n[a] = ctx->n[a];
    d.kresp = ctx->solver_mode == STZ_SOLVER_LOCAL ? ctx->kresp : nullptr;
    d.local_resync = ctx->local_resync;
    d.local_max = std::min(ctx->local_max, STZ_LOCAL_MAXSRC);
}

// host scalars -> device run state; positions the lookup window at the stream position
int push_state(stz_ctx *ctx, long long budget)
{
    StzResident *rs = ctx->rs;
    // the stream position the next step consumes from
    long long pos;
    if (rs->rng_lazy && rs->win_ok && rs->win_epoch == ctx->rng_epoch) {
        pos = rs->rng_target;
    } else {
        rs->rng_lazy = false;
        pos = ctx->rng.count;
    }
    const StzMtWindow &W = rs->win;
    const bool inside = rs->win_ok && rs->win_epoch == ctx->rng_epoch && pos >= W.base && pos <= W.base + W.len;
    if (!inside || rs->force_rebuild || W.base + W.len - pos < rs->wsize / 4) {
        rs->force_rebuild = false;
        StzRng from;
        if (inside) {
            from = W.at_base;
            from.skip(pos - W.base);
        } else {
            from = ctx->rng;          // not lazy here: ctx->rng is at pos
        }
        int rc = window_build(ctx, from);
        if (rc) return rc;
    }
    RepState &s = rs->hst;
    std::memset(&s, 0, sizeof(s));
    for (int c = 0; c < 6; ++c) {
        s.epsavg[c] = ctx->epsavg[c];
        s.stress[c] = ctx->stress[c];
        s.soff[c] = ctx->soff[c];
    }
    s.totime = ctx->totime;
    s.pend_comp = ctx->pend_comp;
    s.pend_delta = ctx->pend_delta;
    s.strainhist_last = ctx->strainhist_last;
    s.soc_elast = ctx->soc_elast;
    s.endstrain = ctx->endstrain;
    s.printperiod = ctx->printperiod;
    s.print_on = ctx->print_on;
    s.step = ctx->step;
    s.kstep = ctx->kstep;
    s.astep = ctx->astep;
    s.estep = ctx->estep;
    s.ntrans = ctx->ntrans;
    s.nsolve = ctx->nsolve;
    s.avlan = ctx->avlan;
    for (int c = 0; c < 6; ++c) {
        s.ebase[c] = ctx->ebase[c];
        s.usum[c] = ctx->usum[c];
    }
    s.nlocal = ctx->nlocal;
    s.rpos = pos;
    s.budget = std::min<long long>(budget, rs->rec_cap);
    s.halt = budget > 0 ? HALT_NONE : HALT_BUDGET;
    s.nrec = 0;
    RS_HIP(ctx, hipMemcpyAsync(rs->st, &s, sizeof(s), hipMemcpyHostToDevice, ctx->stream));
    __atomic_store_n(rs->hflag, s.halt, __ATOMIC_RELAXED);
    const int none[2] = {KIND_NONE, 0};
    RS_HIP(ctx, hipMemcpyAsync(ctx->res + RES_KIND, &none[0], sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RS_HIP(ctx, hipMemcpyAsync(ctx->res + RES_FFT, &none[0], sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RS_HIP(ctx, hipMemcpyAsync(ctx->res + RES_LOCAL, &none[1], sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    fill_desc(ctx);
    RS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return STZ_OK;
}

// device run state -> host scalars (rs->hst must hold the downloaded state)
void pull_state(stz_ctx *ctx)
{
    StzResident *rs = ctx->rs;
    const RepState &s = rs->hst;
    for (int c = 0; c < 6; ++c) {
        ctx->epsavg[c] = s.epsavg[c];
        ctx->stress[c] = s.stress[c];
        ctx->soff[c] = s.soff[c];
    }
    ctx->totime = s.totime;
    ctx->pend_comp = s.pend_comp;
    ctx->pend_delta = s.pend_delta;
    ctx->strainhist_last = s.strainhist_last;
    ctx->soc_elast = s.soc_elast;
    ctx->step = s.step;
    ctx->kstep = s.kstep;
    ctx->astep = s.astep;
    ctx->estep = s.estep;
    ctx->ntrans = s.ntrans;
    ctx->nsolve = s.nsolve;
    ctx->avlan = s.avlan;
    for (int c = 0; c < 6; ++c) {
        ctx->ebase[c] = s.ebase[c];
        ctx->usum[c] = s.usum[c];
    }
    ctx->nlocal = s.nlocal;
    if (!(rs->rng_lazy == false && ctx->rng.count == s.rpos)) {
        rs->rng_lazy = true;
        rs->rng_target = s.rpos;
    }
}

// --------------------------------------------------------------------------------
// plans
// --------------------------------------------------------------------------------
template <typename R, int DIM>
int launch_class(StzPlan *pl, StzClass &k, hipEvent_t ea, hipEvent_t eb)
{
    stz_ctx *p = k.reps[0];
    const int nrep = (int)k.reps.size();
    hipStream_t s = k.stream ? k.stream : pl->stream;
    if (ea) (void)hipEventRecord(ea, s);
    hipLaunchKernelGGL((kb_rates<R, DIM>), dim3((unsigned)p->nblk, nrep), dim3(NT), 0, s, k.d_reps, p->N, p->nblk, p->M);
    if (eb) (void)hipEventRecord(eb, s);
    const unsigned gb = (unsigned)((p->ngrp + STZ_GRP_PER_BLOCK - 1) / STZ_GRP_PER_BLOCK);
    hipLaunchKernelGGL(kb_group, dim3(gb, nrep), dim3(NT), 0, s, k.d_reps, (long long)p->M * p->nblk, p->ngrp);
    const long long nch = std::max((p->ngrp + NT - 1) / NT, (p->nblk + NT - 1) / NT);
    hipLaunchKernelGGL(kb_chunks, dim3((unsigned)nch, nrep), dim3(NT), 0, s, k.d_reps, p->ngrp, p->nblk);
    hipLaunchKernelGGL((kb_decide<R, DIM>), dim3(nrep), dim3(NT), 0, s, k.d_reps, p->N, p->nblk, p->ngrp, p->M);
    hipLaunchKernelGGL((kb_athermal<R, DIM>), dim3((unsigned)((p->nblk + NT - 1) / NT), nrep), dim3(NT), 0,
                       s, k.d_reps, p->N, p->M, p->nblk);
    if (hipGetLastError() != hipSuccess) return STZ_E_HIP;
    int nparts = 0;
    if (k.fft) {            // replicas whose every transform is followed by the full solve
        const int saved = p->timing;
        p->timing = 0;
        const int rc = stz_solve_launch_batched(p, nrep, k.d_solve, s, &nparts);
        p->timing = saved;
        if (rc) return rc;
    }
    const int fp64 = p->prec == STZ_FP64;
    const long long nlb = stz_local_blocks(p);
    if (k.local)
        hipLaunchKernelGGL((kb_local<R, DIM>), dim3((unsigned)nlb, nrep), dim3(NT_LOCAL), 0, s, k.d_reps, fp64);
    hipLaunchKernelGGL((kb_finish<DIM>), dim3(nrep), dim3(STZ_NT_MEAN), 0, s, k.d_reps, fp64, nparts, (int)nlb,
                       1.0 / (double)p->N, p->nblk);
    return hipGetLastError() == hipSuccess ? STZ_OK : STZ_E_HIP;
}

// One step of every class.  With several classes each runs on its own stream between a
// fork and a join on the plan's stream: the classes share no memory, and the short
// latency-bound kernels of one class overlap the bandwidth-bound passes of another
// (a captured graph holds them as parallel branches).
int launch_step(StzPlan *pl, hipEvent_t ea = nullptr, hipEvent_t eb = nullptr)
{
    if (pl->pactive) {                  // the whole step of every class in one launch
        const stz_ctx *p = pl->cls[0].reps[0];
        const dim3 g(pl->hargs.total), b(NT);       // one workgroup per ticket
        if (p->prec == STZ_FP64) {
            if (p->dim == 2) hipLaunchKernelGGL((kp_step<double, 2>), g, b, 0, pl->stream, pl->pargs);
            else hipLaunchKernelGGL((kp_step<double, 3>), g, b, 0, pl->stream, pl->pargs);
        } else {
            if (p->dim == 2) hipLaunchKernelGGL((kp_step<float, 2>), g, b, 0, pl->stream, pl->pargs);
            else hipLaunchKernelGGL((kp_step<float, 3>), g, b, 0, pl->stream, pl->pargs);
        }
        return hipGetLastError() == hipSuccess ? STZ_OK : STZ_E_HIP;
    }
    if (pl->fork) {
        if (hipEventRecord(pl->fork, pl->stream) != hipSuccess) return STZ_E_HIP;
        for (StzClass &k : pl->cls)
            if (hipStreamWaitEvent(k.stream, pl->fork, 0) != hipSuccess) return STZ_E_HIP;
    }
    for (StzClass &k : pl->cls) {
        stz_ctx *p = k.reps[0];
        int rc;
        if (p->prec == STZ_FP64)
            rc = p->dim == 2 ? launch_class<double, 2>(pl, k, ea, eb) : launch_class<double, 3>(pl, k, ea, eb);
        else
            rc = p->dim == 2 ? launch_class<float, 2>(pl, k, ea, eb) : launch_class<float, 3>(pl, k, ea, eb);
        if (rc) return rc;
    }
    if (pl->fork) {
        for (StzClass &k : pl->cls) {
            if (hipEventRecord(k.done, k.stream) != hipSuccess) return STZ_E_HIP;
            if (hipStreamWaitEvent(pl->stream, k.done, 0) != hipSuccess) return STZ_E_HIP;
        }
    }
    return STZ_OK;
}

void plan_free(StzPlan *pl)
{
    if (!pl) return;
    for (hipEvent_t e : pl->eva) (void)hipEventDestroy(e);
    for (hipEvent_t e : pl->evc) (void)hipEventDestroy(e);
    for (hipEvent_t e : pl->evb) (void)hipEventDestroy(e);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    if (pl->psync) (void)hipFree(pl->psync);
    if (pl->pargs) (void)hipFree(pl->pargs);
    for (StzClass &k : pl->cls) {
        if (k.stream) (void)hipStreamSynchronize(k.stream);
        if (k.d_reps) (void)hipFree(k.d_reps);
        if (k.d_solve) (void)hipFree(k.d_solve);
        if (k.done) (void)hipEventDestroy(k.done);
        if (k.stream) (void)hipStreamDestroy(k.stream);
    }
    if (pl->fork) (void)hipEventDestroy(pl->fork);
    delete pl;
}

// groups the contexts into classes of one shape, precision and elasticity
int plan_make(stz_ctx *const *ctxs, int n, StzPlan **out)
{
    StzPlan *pl = new StzPlan();
    pl->owner = ctxs[0];
    pl->stream = ctxs[0]->stream;
    std::map<std::tuple<int, int, int, int, int, int, double, double, int>, int> key2cls;
    for (int i = 0; i < n; ++i) {
        stz_ctx *c = ctxs[i];
        int rc = ensure_resident(c);
        if (rc) {
            plan_free(pl);
            return rc;
        }
        auto key = std::make_tuple(c->dim, c->n[0], c->n[1], c->n[2], c->M, c->prec, c->k.lambda, c->k.mu, c->p.loading);
        auto it = key2cls.find(key);
        if (it == key2cls.end()) {
            key2cls[key] = (int)pl->cls.size();
            pl->cls.emplace_back();
            it = key2cls.find(key);
        }
        pl->cls[(size_t)it->second].reps.push_back(c);
        pl->all.push_back(c);
    }
    for (StzClass &k : pl->cls) {
        const size_t nr = k.reps.size();
        if (hipMalloc(&k.d_reps, nr * sizeof(RepDev)) != hipSuccess || hipMalloc(&k.d_solve, nr * sizeof(SolveRep)) != hipSuccess) {
            plan_free(pl);
            return STZ_E_HIP;
        }
    }
    if (pl->cls.size() > 1) {
        bool ok = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming) == hipSuccess;
        for (StzClass &k : pl->cls) {
            ok = ok && hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking) == hipSuccess;
            ok = ok && hipEventCreateWithFlags(&k.done, hipEventDisableTiming) == hipSuccess;
        }
        if (!ok) {
            plan_free(pl);
            return STZ_E_HIP;
        }
    }
    *out = pl;
    return STZ_OK;
}

// why a plan cannot take the persistent step, or nullptr
const char *persist_refusal(const std::vector<stz_ctx *> &all, size_t nclasses)
{
    if (nclasses > (size_t)PMAXCLS) return "the persistent step takes at most 8 shape classes";
    for (const stz_ctx *c : all) {
        if (c->prec != all[0]->prec || c->dim != all[0]->dim)
            return "the persistent step takes replicas of one precision and dimension";
        if (c->solver_mode != STZ_SOLVER_LOCAL)
            return "the persistent step takes replicas on the event-local stress update";
    }
    return nullptr;
}

// The launch description of the persistent step (tickets, phases, classes); pactive
// stays false when the plan's replicas do not allow it.  Items per ticket: 4 rate blocks
// (a ticket and a count update per 4 x 256 voxels), 8 athermal chunks (most return at
// once), 2 local-update row groups, one item of the other phases.
int plan_persist_setup(StzPlan *pl)
{
    pl->pactive = false;
    if (!pl->persist || persist_refusal(pl->all, pl->cls.size())) return STZ_OK;
    for (const StzClass &k : pl->cls)
        if (k.fft || !k.local) return STZ_OK;
    stz_ctx *o = pl->owner;
    std::vector<int> order(pl->cls.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pl->cls[(size_t)a].reps[0]->N > pl->cls[(size_t)b].reps[0]->N; });
    std::vector<PShape> shp;
    pl->hreps.clear();
    for (int j : order) {
        const StzClass &k = pl->cls[(size_t)j];
        const stz_ctx *p = k.reps[0];
        for (const stz_ctx *c : k.reps) pl->hreps.push_back(c->rs->hdev);
        shp.push_back(PShape{p->N, p->nblk, p->ngrp, stz_local_blocks(p), p->M, (int)k.reps.size()});
    }
    PArgs &A = pl->hargs;
    if (!persist_table(shp.data(), (int)shp.size(), A)) return STZ_OK;
    const int nrep = A.nrep;
    A.fp64 = o->prec == STZ_FP64;
    if (pl->psync_n < psync_words(nrep)) {
        if (pl->psync) (void)hipFree(pl->psync);
        pl->psync = nullptr;
        pl->psync_n = 0;
        RS_HIP(o, hipMalloc((void **)&pl->psync, sizeof(unsigned) * (size_t)psync_words(nrep)));
        RS_HIP(o, hipMemset(pl->psync, 0, sizeof(unsigned) * (size_t)psync_words(nrep)));
        pl->psync_n = psync_words(nrep);
    }
    A.sync = pl->psync;
    int dev = 0, khz = 0;
    RS_HIP(o, hipGetDevice(&dev));
    RS_HIP(o, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    A.spin_ticks = (unsigned long long)std::max(khz, 1000) * 1000ull * 5ull;   // 5 s
    if (pl->pargs_rep < nrep) {
        if (pl->pargs) (void)hipFree(pl->pargs);
        pl->pargs = nullptr;
        pl->pargs_rep = 0;
        RS_HIP(o, hipMalloc((void **)&pl->pargs, sizeof(PArgs) + sizeof(RepDev) * (size_t)nrep));
        pl->pargs_rep = nrep;
    }
    RS_HIP(o, hipMemcpyAsync(pl->pargs, &A, sizeof(PArgs), hipMemcpyHostToDevice, pl->stream));
    RS_HIP(o, hipMemcpyAsync(pl->pargs + 1, pl->hreps.data(), sizeof(RepDev) * (size_t)nrep, hipMemcpyHostToDevice,
                             pl->stream));
    pl->pactive = true;
    return STZ_OK;
}

// uploads the replica descriptors (they change when a lookup window is rebuilt)
int plan_upload(StzPlan *pl)
{
    stz_ctx *o = pl->owner;
    for (StzClass &k : pl->cls) {
        std::vector<RepDev> d;
        std::vector<SolveRep> s;
        k.local = false;
        k.fft = false;
        for (stz_ctx *c : k.reps) {
            d.push_back(c->rs->hdev);
            k.local = k.local || c->rs->hdev.kresp != nullptr;
            k.fft = k.fft || c->rs->hdev.kresp == nullptr;
            SolveRep q;
            q.eps0 = c->eps0;
            q.spec = c->spec;
            q.sigma = c->sigma;
            q.spart = c->spart;
            q.res = c->res;
            q.epsavg = c->rs->st->epsavg;
            s.push_back(q);
        }
        RS_HIP(o, hipMemcpyAsync(k.d_reps, d.data(), d.size() * sizeof(RepDev), hipMemcpyHostToDevice, pl->stream));
        RS_HIP(o, hipMemcpyAsync(k.d_solve, s.data(), s.size() * sizeof(SolveRep), hipMemcpyHostToDevice, pl->stream));
    }
    const bool was = pl->pactive;
    int rc = plan_persist_setup(pl);
    if (rc) return rc;
    if (was != pl->pactive) pl->graph_epoch = -1;   // a captured graph holds the other launch sequence
    RS_HIP(o, hipStreamSynchronize(pl->stream));
    return STZ_OK;
}

// captures use_graph consecutive steps into one graph
int plan_graph(StzPlan *pl)
{
    long long ep = 0;
    for (const stz_ctx *c : pl->all) ep += c->tune_epoch;
    if (pl->exec && pl->graph_steps == pl->use_graph && pl->graph_epoch == ep) return STZ_OK;
    pl->graph_epoch = ep;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    pl->exec = nullptr;
    pl->graph = nullptr;
    stz_ctx *o = pl->owner;
    RS_HIP(o, hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
    int rc = STZ_OK;
    for (int i = 0; i < pl->use_graph && !rc; ++i) rc = launch_step(pl);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(pl->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        o->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return STZ_E_HIP;
    }
    pl->graph = g;
    RS_HIP(o, hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0));
    pl->graph_steps = pl->use_graph;
    return STZ_OK;
}

// true when every replica of the plan has halted (the device publishes halts to mapped
// host memory, so the host stops enqueueing steps that would only run as no-ops)
bool all_halted(const StzPlan *pl)
{
    for (const stz_ctx *c : pl->all)
        if (__atomic_load_n(c->rs->hflag, __ATOMIC_RELAXED) == HALT_NONE) return false;
    return true;
}

// Enqueues up to k steps in chunks, keeping at most `depth` steps ahead of the GPU,
// and stops once every replica has halted.  Steps beyond a replica's budget or after
// its halt run as no-ops.
int plan_launch_steps(StzPlan *pl, int k)
{
    stz_ctx *o = pl->owner;
    const int chunk = pl->use_graph > 0 ? pl->use_graph : 2;
    const int depth = std::max(3 * chunk, 6);
    if (pl->use_graph > 0) {
        int rc = plan_graph(pl);
        if (rc) return rc;
    }
    stz_ctx *p = pl->all[0];
    const bool timed = pl->use_graph == 0 && pl->all.size() == 1 && p->timing && !pl->pactive;
    if (timed && pl->eva.size() < (size_t)k) {
        while (pl->eva.size() < (size_t)k) {
            hipEvent_t a, b;
            RS_HIP(o, hipEventCreate(&a));
            RS_HIP(o, hipEventCreate(&b));
            pl->eva.push_back(a);
            pl->evb.push_back(b);
        }
    }
    const int nchunks = (k + chunk - 1) / chunk;
    while ((int)pl->evc.size() < nchunks) {
        hipEvent_t e;
        RS_HIP(o, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        pl->evc.push_back(e);
    }
    int launched = 0, waited = 0, nc = 0;
    while (launched < k) {
        const int c = std::min(chunk, k - launched);
        if (pl->use_graph > 0) {
            RS_HIP(o, hipGraphLaunch(pl->exec, pl->stream));
        } else {
            for (int i = 0; i < c; ++i) {
                const int j = launched + i;
                int rc = launch_step(pl, timed ? pl->eva[(size_t)j] : nullptr, timed ? pl->evb[(size_t)j] : nullptr);
                if (rc) return rc;
            }
        }
        launched += c;
        RS_HIP(o, hipEventRecord(pl->evc[(size_t)nc++], pl->stream));
        while ((launched - waited) > depth) {
            RS_HIP(o, hipEventSynchronize(pl->evc[(size_t)(waited / chunk)]));
            waited += chunk;
        }
        if (all_halted(pl)) break;
    }
    pl->timed = timed ? launched : 0;
    return STZ_OK;
}

}  // namespace

// re-positions ctx->rng at the resident stream position when a resident run moved it
void stz_rng_sync(stz_ctx *ctx)
{
    StzResident *rs = ctx->rs;
    if (!rs || !rs->rng_lazy) return;
    StzRng g = rs->win.at_base;
    g.skip(rs->rng_target - rs->win.base);
    ctx->rng = g;
    rs->rng_lazy = false;
}

void stz_rng_replaced(stz_ctx *ctx)
{
    ctx->rng_epoch++;
    if (ctx->rs) ctx->rs->rng_lazy = false;
}

void stz_resident_free(stz_ctx *ctx)
{
    StzResident *rs = ctx->rs;
    if (!rs) return;
    plan_free(rs->self);
    free_window(rs);
    if (rs->st) (void)hipFree(rs->st);
    if (rs->rec) (void)hipFree(rs->rec);
    if (rs->hflag) (void)hipHostFree(rs->hflag);
    delete rs;
    ctx->rs = nullptr;
}

// Runs up to nsteps resident steps on every context of the plan.  results[i] (optional)
// receives up to nsteps records of context i, done[i] their count, rcs[i] its return
// code (0 running, 1 terminated, < 0 error).  A print step is re-solved on the host
// (STZ/bmg.c:289-295); with stop_at_print the context stops right after the print step,
// before that re-solve, which the caller then runs with stz_step_print.
int stz_plan_run(StzPlan *pl, int nsteps, stz_step_result *const *results, int *done, int *rcs, int stop_at_print)
{
    const size_t n = pl->all.size();
    std::vector<int> active(n, 1);
    for (size_t i = 0; i < n; ++i) {
        done[i] = 0;
        rcs[i] = 0;
    }
    for (size_t i = 0; i < n; ++i) {
        stz_ctx *c = pl->all[i];
        int rc = push_state(c, nsteps);
        if (rc) return rc;
    }
    int rc = plan_upload(pl);
    if (rc) return rc;
    for (;;) {
        long long most = 0;
        for (size_t i = 0; i < n; ++i)
            if (active[i]) most = std::max<long long>(most, nsteps - done[i]);
        if (most <= 0) break;
        const int k = (int)std::min<long long>(most, pl->batch);
        rc = plan_launch_steps(pl, k);
        if (rc) return rc;
        stz_ctx *o = pl->owner;
        for (size_t i = 0; i < n; ++i) {
            stz_ctx *c = pl->all[i];
            RS_HIP(o, hipMemcpyAsync(&c->rs->hst, c->rs->st, sizeof(RepState), hipMemcpyDeviceToHost, pl->stream));
        }
        unsigned aborted = 0;
        if (pl->pactive)
            RS_HIP(o, hipMemcpyAsync(&aborted, pl->psync + PSYNC_ABORT, sizeof(unsigned), hipMemcpyDeviceToHost, pl->stream));
        RS_HIP(o, hipStreamSynchronize(pl->stream));
        if (aborted) {
            // a dependency wait of the persistent step ran out of time: the replicas' state
            // is not trustworthy any more
            RS_HIP(o, hipMemset(pl->psync, 0, sizeof(unsigned) * (size_t)pl->psync_n));
            for (size_t i = 0; i < n; ++i) {
                pl->all[i]->err = "persistent step: a dependency wait exceeded its time limit";
                rcs[i] = STZ_E_HIP;
            }
            return STZ_E_HIP;
        }
        if (pl->timed > 0) {
            // only the steps that recorded ran the rate pass: the launches after a halt
            // (print, budget, window, error) return at once and are left out of the mean
            stz_ctx *p = pl->all[0];
            const int worked = std::min(pl->timed, p->rs->hst.nrec);
            for (int i = 0; i < worked; ++i) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, pl->eva[(size_t)i], pl->evb[(size_t)i]) == hipSuccess) {
                    p->tms[0] += ms;
                    p->tn[0] += 1;
                }
            }
            pl->timed = 0;
        }
        bool need_upload = false;
        for (size_t i = 0; i < n; ++i) {
            if (!active[i]) continue;
            stz_ctx *c = pl->all[i];
            StzResident *rs = c->rs;
            const int nrec = rs->hst.nrec;
            if (nrec > 0) {
                RS_HIP(o, hipMemcpy(rs->hrec.data(), rs->rec, (size_t)nrec * sizeof(stz_step_result), hipMemcpyDeviceToHost));
                if (results && results[i])
                    std::memcpy(results[i] + done[i], rs->hrec.data(), (size_t)nrec * sizeof(stz_step_result));
                done[i] += nrec;
            }
            const int halt = rs->hst.halt;
            pull_state(c);
            long long next_budget = nsteps - done[i];
            if (halt == HALT_ERR) {
                c->err = rs->hst.err == STZ_E_SELECT ? "KMC failed to locate an activation mode"
                         : rs->hst.err == STZ_E_NAN  ? "NaN in the stress solve"
                                                     : "resident step: unexpected step kind";
                rcs[i] = rs->hst.err;
                active[i] = 0;
                continue;
            }
            if (halt == HALT_DONE) {
                rcs[i] = 1;
                active[i] = 0;
                continue;
            }
            if (halt == HALT_PRINT && stop_at_print) {
                // the caller writes the print block and then calls stz_step_print
                active[i] = 0;
                continue;
            }
            if (halt == HALT_PRINT) {
                // print-step re-solve on the host path, STZ/bmg.c:289-295
                stz_step_result tmp;
                int prc = stz_step_print(c, &tmp);
                if (prc < 0) {
                    rcs[i] = prc;
                    active[i] = 0;
                    continue;
                }
                if (results && results[i] && nrec > 0)
                    for (int cc = 0; cc < 6; ++cc) results[i][done[i] - 1].stress[cc] = c->stress[cc];
                if (prc == 1) {
                    rcs[i] = 1;
                    active[i] = 0;
                    continue;
                }
            }
            if (halt == HALT_SOLVE) {
                // the step needs the full solve (periodic re-solve of the local update, or an
                // avalanche larger than it takes): it runs on the host path, whose stress
                // solve is the full FFT solve, with the same random numbers
                stz_rng_sync(c);
                stz_step_result r;
                int hrc = stz_step_kmc(c, &r);
                if (hrc >= 0 && r.printed && !stop_at_print) hrc = stz_step_print(c, &r);
                if (hrc >= 0) {
                    if (results && results[i]) results[i][done[i]] = r;
                    done[i] += 1;
                    next_budget -= 1;
                }
                if (hrc != 0) {
                    rcs[i] = hrc;
                    active[i] = 0;
                    continue;
                }
                if (r.printed && stop_at_print) {
                    active[i] = 0;
                    continue;
                }
            }
            if (halt == HALT_RNG) {
                // the window does not hold the coming step's draws
                const long long pos = rs->hst.rpos;
                const bool fresh = rs->win.base == pos;
                if (fresh && rs->win.len == rs->wsize && rs->wsize < (1ll << 28)) {
                    rs->wsize *= 4;           // a larger avalanche than the window holds
                } else if (fresh) {
                    // a zero word ahead (or the largest window is too small): one step on the
                    // host path, which draws from the stream itself
                    stz_rng_sync(c);
                    stz_step_result r;
                    int hrc = stz_step(c, &r);
                    if (results && results[i]) results[i][done[i]] = r;
                    done[i] += 1;
                    next_budget -= 1;
                    if (hrc != 0) {
                        rcs[i] = hrc;
                        active[i] = 0;
                        continue;
                    }
                }
                rs->force_rebuild = true;
            }
            if (next_budget <= 0) {
                active[i] = 0;
                continue;
            }
            rc = push_state(c, next_budget);
            if (rc) return rc;
            need_upload = true;
        }
        if (need_upload) {
            rc = plan_upload(pl);
            if (rc) return rc;
        }
    }
    return STZ_OK;
}

int stz_plan_create(stz_ctx *const *ctxs, int n, StzPlan **out) { return plan_make(ctxs, n, out); }
void stz_plan_destroy(StzPlan *pl) { plan_free(pl); }
void stz_plan_config(StzPlan *pl, int batch, int use_graph)
{
    if (batch > 0) pl->batch = batch;
    if (use_graph >= 0 && use_graph != pl->use_graph) {
        pl->use_graph = use_graph;
    }
}
StzPlan *stz_self_plan(stz_ctx *ctx)
{
    if (ensure_resident(ctx)) return nullptr;
    if (!ctx->rs->self) {
        StzPlan *pl = nullptr;
        if (plan_make(&ctx, 1, &pl)) return nullptr;
        ctx->rs->self = pl;
    }
    return ctx->rs->self;
}

// ==================================================================================
// C-ABI
// ==================================================================================
struct stz_ensemble {
    StzPlan *pl = nullptr;
};

namespace {

int set_batch(stz_ctx *ctx, int batch)
{
    int rc = ensure_resident(ctx);
    if (rc) return rc;
    StzResident *rs = ctx->rs;
    if (batch > rs->rec_cap) {
        if (rs->rec) (void)hipFree(rs->rec);
        rs->rec = nullptr;
        rs->rec_cap = 0;
        rc = rs_alloc(ctx, (void **)&rs->rec, sizeof(stz_step_result) * (size_t)batch);
        if (rc) return rc;
        rs->rec_cap = batch;
        rs->hrec.resize((size_t)batch);
    }
    rs->batch = batch;
    return STZ_OK;
}

}  // namespace

extern "C" {

int stz_run_config(stz_ctx *ctx, int batch, int use_graph, long long window_words)
{
    if (!ctx || batch < 0 || batch > 65536 || window_words < 0 || use_graph > 1024) return STZ_E_ARG;
    StzPlan *pl = stz_self_plan(ctx);
    if (!pl) return STZ_E_HIP;
    if (batch > 0) {
        int rc = set_batch(ctx, batch);
        if (rc) return rc;
        pl->batch = batch;
    }
    if (use_graph >= 0) pl->use_graph = use_graph;
    if (window_words > 0) {
        if (window_words < 4096 || window_words > (1ll << 28)) return STZ_E_ARG;
        ctx->rs->wsize = window_words;
        ctx->rs->force_rebuild = true;
    }
    return STZ_OK;
}

int stz_run_steps(stz_ctx *ctx, int nsteps, stz_step_result *results, int *done)
{
    if (!ctx || nsteps < 0) return STZ_E_ARG;
    int d = 0, rc = 0;
    if (ctx->lamx || ctx->heat || ctx->nmro) {
        // per-voxel moduli (the fixed-point loop ends on a host decision each iteration), heat
        // conduction, or crystal-like MRO voxels (their transforms draw uniforms instead of
        // Gaussians): every step runs on the host path (same random numbers, same results)
        const bool stop_at_print = ctx->rs && ctx->rs->stop_at_print;
        for (int i = 0; i < nsteps; ++i) {
            stz_step_result r;
            int hrc = stz_step_kmc(ctx, &r);
            if (hrc >= 0 && r.printed && !stop_at_print) hrc = stz_step_print(ctx, &r);
            if (hrc >= 0) {
                if (results) results[d] = r;
                ++d;
            }
            if (hrc != 0) {
                rc = hrc;
                break;
            }
            if (r.printed && stop_at_print) break;
        }
        if (done) *done = d;
        return rc;
    }
    if (nsteps == 0) {                  // prepare: resident resources and the lookup window
        rc = ensure_resident(ctx);
        if (!rc) rc = push_state(ctx, 0);
    } else {
        StzPlan *pl = stz_self_plan(ctx);
        if (!pl) return STZ_E_HIP;
        stz_step_result *const rr[1] = {results};
        int prc = stz_plan_run(pl, nsteps, rr, &d, &rc, ctx->rs->stop_at_print);
        if (prc) rc = prc;
    }
    if (done) *done = d;
    return rc;
}

int stz_run_persistent(stz_ctx *ctx, int on)
{
    if (!ctx) return STZ_E_ARG;
    StzPlan *pl = stz_self_plan(ctx);
    if (!pl) return STZ_E_HIP;
    if (on) {
        if (const char *why = persist_refusal(pl->all, pl->cls.size())) {
            ctx->err = why;
            return STZ_E_UNSUPPORTED;
        }
    }
    pl->persist = on != 0;
    pl->graph_epoch = -1;
    return STZ_OK;
}

int stz_run_stop_at_print(stz_ctx *ctx, int on)
{
    if (!ctx) return STZ_E_ARG;
    int rc = ensure_resident(ctx);
    if (rc) return rc;
    ctx->rs->stop_at_print = on != 0;
    return STZ_OK;
}

int stz_ensemble_create(stz_ctx *const *members, int n, stz_ensemble **out)
{
    if (!members || n < 1 || !out) return STZ_E_ARG;
    for (int i = 0; i < n; ++i)
        if (!members[i]) return STZ_E_ARG;
    for (int i = 0; i < n; ++i)
        if (members[i]->lamx || members[i]->heat || members[i]->nmro) {
            members[i]->err = "an ensemble steps glass replicas of uniform moduli without heat conduction only";
            return STZ_E_UNSUPPORTED;
        }
    *out = nullptr;
    stz_ensemble *e = new stz_ensemble();
    int rc = stz_plan_create(members, n, &e->pl);
    if (rc) {
        delete e;
        return rc;
    }
    for (int i = 0; i < n; ++i) {
        rc = set_batch(members[i], e->pl->batch);
        if (rc) {
            stz_plan_destroy(e->pl);
            delete e;
            return rc;
        }
    }
    *out = e;
    return STZ_OK;
}

int stz_ensemble_config(stz_ensemble *e, int batch, int use_graph)
{
    if (!e || batch < 0 || batch > 65536 || use_graph > 1024) return STZ_E_ARG;
    if (batch > 0) {
        for (stz_ctx *c : e->pl->all) {
            int rc = set_batch(c, batch);
            if (rc) return rc;
        }
        e->pl->batch = batch;
    }
    if (use_graph >= 0) e->pl->use_graph = use_graph;
    return STZ_OK;
}

int stz_ensemble_run(stz_ensemble *e, int nsteps, stz_step_result *const *results, int *done, int *rcs)
{
    if (!e || nsteps < 0 || !done || !rcs) return STZ_E_ARG;
    for (stz_ctx *c : e->pl->all)
        if (c->lamx || c->heat || c->nmro) {
            c->err = "an ensemble steps glass replicas of uniform moduli without heat conduction only";
            return STZ_E_UNSUPPORTED;
        }
    return stz_plan_run(e->pl, nsteps, results, done, rcs, 0);
}

// CPU check of the persistent step's ticket table (no device call): builds the table of
// ncls classes (shape[c] = {dim, n0, n1, n2, M, precision, replicas}, in launch order),
// decodes every ticket as the kernel does, and counts violations: a ticket outside its
// segment's ranges, an item covered other than exactly once, or a ticket that a ticket of
// an earlier phase of its replica follows.  out = {tickets, items, violations}.
int stz_persist_probe(int ncls, const int *shape, long long *out)
{
    if (ncls < 1 || ncls > PMAXCLS || !shape || !out) return STZ_E_ARG;
    std::vector<PShape> shp;
    for (int c = 0; c < ncls; ++c) {
        const int *q = shape + 7 * c;
        const int dim = q[0], n[3] = {q[1], q[2], dim == 2 ? 1 : q[3]}, M = q[4], prec = q[5], nrep = q[6];
        if ((dim != 2 && dim != 3) || n[0] < 1 || n[1] < 1 || n[2] < 1 || M < 1 || M > STZ_MAXM || nrep < 1)
            return STZ_E_ARG;
        PShape p;
        p.N = (long long)n[0] * n[1] * n[2];
        p.nblk = (p.N + STZ_BLK - 1) / STZ_BLK;
        p.ngrp = ((long long)M * p.nblk + STZ_GRP - 1) / STZ_GRP;
        const LocalGeom g = local_geom(dim, n, local_row_voxels(prec == STZ_MIXED));
        p.nlb = (g.items + g.ipw - 1) / g.ipw;
        p.M = M;
        p.nrep = nrep;
        shp.push_back(p);
    }
    PArgs A;
    if (!persist_table(shp.data(), ncls, A)) return STZ_E_TOO_LARGE;
    long long bad = 0, items = 0;
    // coverage per (replica, phase, item), first/last ticket per (replica, phase)
    std::vector<std::vector<std::vector<unsigned char>>> cov((size_t)A.nrep);
    std::vector<long long> first((size_t)A.nrep * NPH, -1), last((size_t)A.nrep * NPH, -1);
    for (int c = 0; c < ncls; ++c)
        for (int r = 0; r < shp[(size_t)c].nrep; ++r) {
            auto &v = cov[(size_t)(A.cls[c].rep0 + r)];
            v.resize(NPH);
            for (int ph = 0; ph < NPH; ++ph) {
                v[(size_t)ph].assign((size_t)A.cls[c].items[ph], 0);
                items += A.cls[c].items[ph];
            }
        }
    for (unsigned t = 0; t < A.total; ++t) {
        int ph, ci, r;
        long long i0, i1;
        persist_decode(&A, t, ph, ci, r, i0, i1);
        if (ph < 0 || ph >= NPH || ci < 0 || ci >= ncls || r < 0 || r >= shp[(size_t)ci].nrep || i0 < 0 || i0 >= i1 ||
            i1 > A.cls[ci].items[ph]) {
            ++bad;
            continue;
        }
        const int gr = A.cls[ci].rep0 + r;
        for (long long i = i0; i < i1; ++i) cov[(size_t)gr][(size_t)ph][(size_t)i]++;
        long long &f = first[(size_t)gr * NPH + ph], &l = last[(size_t)gr * NPH + ph];
        if (f < 0) f = t;
        l = t;
    }
    for (int gr = 0; gr < A.nrep; ++gr)
        for (int ph = 0; ph < NPH; ++ph) {
            for (unsigned char x : cov[(size_t)gr][(size_t)ph]) bad += x != 1;
            // every ticket of an earlier phase comes first (the waits only look back)
            for (int q = 0; q < ph; ++q)
                if (last[(size_t)gr * NPH + q] >= 0 && first[(size_t)gr * NPH + ph] >= 0 &&
                    last[(size_t)gr * NPH + q] > first[(size_t)gr * NPH + ph])
                    ++bad;
        }
    out[0] = A.total;
    out[1] = items;
    out[2] = bad;
    return STZ_OK;
}

int stz_ensemble_persistent(stz_ensemble *e, int on)
{
    if (!e) return STZ_E_ARG;
    StzPlan *pl = e->pl;
    if (on) {
        if (const char *why = persist_refusal(pl->all, pl->cls.size())) {
            pl->owner->err = why;
            return STZ_E_UNSUPPORTED;
        }
    }
    pl->persist = on != 0;
    pl->graph_epoch = -1;
    return STZ_OK;
}

void stz_ensemble_destroy(stz_ensemble *e)
{
    if (!e) return;
    if (e->pl) {
        (void)hipStreamSynchronize(e->pl->stream);
        stz_plan_destroy(e->pl);
    }
    delete e;
}

}  // extern "C"

extern "C" {

// Host check of the lookup window against the stream drawn in order: seeds a generator,
// skips `skip` words, builds a window of window_words words there and takes, from the
// word `offset` into it, either a uniform and then ngauss gaussians (with_uniform = 1)
// or ngauss gaussians, both from the window (with the kernels' index arithmetic) and
// by drawing.  Outputs the values and the stream positions after them.
int stz_mt_window_probe(unsigned long seed, long long skip, long long window_words, long long offset, int with_uniform,
                        int ngauss, double *win_out, double *seq_out, long long *win_end, long long *seq_end)
{
    if (window_words < 16 || offset < 0 || offset >= window_words || ngauss < 0 || !win_out || !seq_out)
        return STZ_E_ARG;
    StzRng g;
    g.seed(seed);
    g.skip(skip);
    StzMtWindow W;
    W.build(g, window_words, 4);
    MtDev m;
    m.w = W.words.data();
    for (int p = 0; p < 2; ++p) {
        m.g[p] = W.g[p].data();
        m.rank[p] = W.rank[p].data();
        m.pair[p] = W.pair[p].data();
        m.nacc[p] = W.nacc[p];
    }
    m.base = W.base;
    m.len = W.len;
    // from the window
    long long pos = W.base + offset;
    int o = 0;
    if (with_uniform) {
        if (pos - m.base >= m.len) return STZ_E_ARG;
        win_out[o++] = m.w[pos - m.base] / 4294967296.0;
        pos += 1;
    }
    if (ngauss > 0) {
        int p;
        const long long k0 = mt_first(m, pos, p);
        if (k0 + ngauss > m.nacc[p]) return STZ_E_ARG;
        for (int i = 0; i < ngauss; ++i) win_out[o + i] = m.g[p][k0 + i];
        pos = mt_after(m, p, k0 + ngauss - 1);
    }
    if (win_end) *win_end = pos;
    // drawn in order
    StzRng s = g;
    s.skip(offset);
    o = 0;
    if (with_uniform) seq_out[o++] = s.uniform();
    for (int i = 0; i < ngauss; ++i) seq_out[o + i] = s.gaussian(1.0);
    if (seq_end) *seq_end = s.count;
    return STZ_OK;
}

}  // extern "C"
