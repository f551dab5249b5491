// stz_internal.h -- context layout and device helpers shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "stz.h"
#include "stz_rng.h"

#define STZ_MAXC 6
#define STZ_MAXS 5
#define STZ_MAXM 64
#define STZ_BLK 256          // voxels per rate block (one per thread)
#define STZ_GRP 256          // rate blocks per selection group
#define STZ_CHUNK 256        // group sums per chunk of the running scan; voxel blocks per athermal chunk
#define STZ_LOCAL_MAXSRC 64  // sites one event-local stress update takes (larger avalanches: full solve)

// control block (ctx->res, doubles unless noted), shared by kernels and host:
//   [0..5] <sigma>  [8] R  [9] #(Q<0)  [10..12] site, mode, Q  [13] #athermal
//   [15] NaN flag (int)  [16..21] mixed stress offsets  [24] Q<0 counter (u64)
//   [26] step kind (int)  [27] dice*R  [28] dice
#define RES_NNEG 24
#define RES_KIND 26
#define RES_TAU 27
#define RES_DICE 28
#define KIND_ELASTIC 0
#define KIND_THERMAL 1
#define KIND_ATHERMAL 2
#define KIND_FAIL 3
#define KIND_ATH_OVERFLOW 4   // athermal avalanche larger than the pre-drawn redraw buffer
#define KIND_STALL 5          // resident step: the random-number window does not hold the step's draws
#define KIND_NEEDSOLVE 6      // resident step with the local update: a full solve is due (nothing changed)
#define KIND_NONE (-1)        // no step decided (resident step halted)
// resident step: the random-number positions of the decided step
#define RES_K0 30             // first accepted-pair index of the step's gaussians (double)
#define RES_PAR 31            // their parity (int)
#define RES_RPOS 32           // stream position after the step (long long)
// resident step: how the decided step's stress is brought up to date (ints)
#define RES_FFT 34            // step kind when the full FFT solve runs, else KIND_NONE (the solve's gate)
#define RES_LOCAL 35          // 1 when the event-local update runs instead
// inhomogeneous solve: sum of the squared eigenstrain change of the last iteration (double)
#define RES_CONV 40

// halt states of a resident replica
#define HALT_NONE 0
#define HALT_DONE 1           // EndStrain or instability reached
#define HALT_PRINT 2          // print step: the host re-solves and writes output
#define HALT_RNG 3            // the random-number window must be refilled
#define HALT_ERR 4            // error code in RepState::err
#define HALT_BUDGET 5         // the batch's step budget is used up
#define HALT_SOLVE 6          // the next step needs the full solve: the host runs it on its own path

// physical constants of the reference, STZ/bmg.h:68-74
#define STZ_EV_IN_J 1.60217733e-19
#define STZ_J_IN_EV (1 / STZ_EV_IN_J)
#define STZ_EPSILON 1e-17
#define STZ_BOLZ 1.380658e-23
#define STZ_PI_REF 3.14159265359
// heating of a transform and heat conduction, STZ/bmg.h:185-188: thermal_coeff1 = 1/Cp/rho
// (Cp = 420 J/kg-K, rho = 6125 kg/m^3), thermal_diffusivity = 3e-6 m^2/s
#define STZ_THERMAL_COEFF1 (1 / 420.0 / 6125.0)
#define STZ_THERMAL_DIFFUSIVITY 3E-6
// crystal-like MRO (six-fold) softening increments, PermSoftmro and TempSoftmro (STZ/init.c:61-63)
#define STZ_PERM_SOFT_MRO 20.0
#define STZ_TEMP_SOFT_MRO 60.0

#ifndef STZ_NT_LOADS
#define STZ_NT_LOADS 1
#endif
// Operands read once per step are loaded non-temporally, so the streams do not evict what
// the step reuses (bit 0: per-mode operands of the rate pass, bit 1: its per-voxel
// operands, bit 2: the local update's response table and stress field)
template <int BIT, typename T>
__device__ __forceinline__ T ld_stream(const T *p)
{
    if constexpr ((STZ_NT_LOADS & BIT) != 0) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

// Global-memory access through pointers the compiler cannot prove global (those read
// from per-replica descriptors): generic pointers compile to flat instructions, and a
// flat store in flight makes every later LDS read wait for it (both count on lgkmcnt).
#define STZ_GLOBAL __attribute__((address_space(1)))
template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <typename T>
__device__ __forceinline__ T gload(const T *p)
{
    return *(const STZ_GLOBAL T *)p;
}
template <typename R>
__device__ __forceinline__ cpx<R> gload(const cpx<R> *p)
{
    typedef typename Vec2<R>::type V;
    const V v = *(const STZ_GLOBAL V *)p;
    return {v.x, v.y};
}
// ld_stream through a global-address-space view of the pointer
template <int BIT, typename T>
__device__ __forceinline__ T gld_stream(const T *p)
{
    if constexpr ((STZ_NT_LOADS & BIT) != 0) return __builtin_nontemporal_load((const STZ_GLOBAL T *)p);
    else return *(const STZ_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, T a)
{
    *(STZ_GLOBAL T *)p = a;
}
template <typename R>
__device__ __forceinline__ void gstore(cpx<R> *p, cpx<R> a)
{
    typedef typename Vec2<R>::type V;
    const V v = {a.x, a.y};
    *(STZ_GLOBAL V *)p = v;
}

// Workgroup-strided loop over TOT elements (element e = tid + i*NT) in two phases per
// chunk of CH iterations: every load of the chunk is issued before the first use, so a
// thread keeps up to CH loads in flight instead of waiting on each (a loop whose body
// loads and then writes LDS is otherwise compiled load -> wait -> write per element).
template <typename T, int TOT, int NT, int CH = 8, typename LoadF, typename UseF>
__device__ __forceinline__ void staged_loop(LoadF load, UseF use)
{
    constexpr int PER = (TOT + NT - 1) / NT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += CH) {
        T v[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int e = tid + (i0 + i) * NT;
            if (i0 + i < PER && (TOT % NT == 0 || e < TOT)) v[i] = load(e);
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int e = tid + (i0 + i) * NT;
            if (i0 + i < PER && (TOT % NT == 0 || e < TOT)) use(e, v[i]);
        }
    }
}

// Constants derived from stz_params exactly as STZ/init.c:6-75 does.
struct StzConsts {
    double E, nu, mu, lambda, alpha, kT, tTemp, dt, voxel, ratepre, wfac_v;
    double permsoft, tempsoft, permcap, thr, half_gamma0;
};

struct FftTwiddles {
    void *tw[3] = {nullptr, nullptr, nullptr};   // per axis, L complex e^{-2 pi i t / L}
};

// MT19937 lookup window in device memory (see StzMtWindow)
struct MtDev {
    const uint32_t *w;
    const double *g[2];
    const int *rank[2];
    const int *pair[2];
    long long base, len;
    long long nacc[2];
};

// first accepted-pair index of the gaussians that start at stream position pos, and its
// parity p; the stream position after the accepted pair k of parity p
__host__ __device__ inline long long mt_first(const MtDev &t, long long pos, int &p)
{
    const long long r = pos - t.base;
    p = (int)(r & 1);
    return t.rank[p][r >> 1];
}
__host__ __device__ inline long long mt_after(const MtDev &t, int p, long long k)
{
    return t.base + p + 2 * ((long long)t.pair[p][k] + 1);
}

// run state of one replica for the resident step (device memory): the reference's
// globals EpsAvg, Stress, totime, the step counters and the stream position
struct RepState {
    double epsavg[6], stress[6], soff[6];
    double totime, pend_delta, strainhist_last, soc_elast, endstrain;
    long long step, kstep, astep, estep, ntrans, nsolve, avlan;
    long long rpos;
    long long budget;        // steps this batch may record
    int pend_comp, halt, err, nrec, printperiod, print_on;
    // event-local stress update: mean strain of the last solve, uniform increments put in
    // the field since (FP64), updates since the last full solve
    double ebase[6], usum[6];
    long long nlocal;
};

// everything the resident kernels need for one replica (device memory)
struct RepDev {
    double *eps0;
    void *sigma, *q0, *deps, *deps_last, *q0eff, *scache, *minq, *spec;
    double *perm, *temp, *tlast;
    double *tloc;
    uint8_t *argm;
    double *bsum, *gsum, *csum, *socp, *spart, *res;
    int *acnt, *ccnt, *aoff, *alist;
    RepState *st;
    stz_step_result *rec;
    int *hflag;              // host-visible copy of the halt state (mapped pinned memory)
    MtDev mt;
    // constants (STZ/init.c:6-75)
    double kT, tTemp, ratepre, lnpre, voxel, thr, dt, permsoft, tempsoft, permcap, half_gamma0;
    double E, nu, mu, lambda, strainrate, strainincrement;
    int loading, lc, athermal_correction, softening;
    // event-local stress update (kresp == nullptr: always the full solve)
    int n[3];
    const void *kresp;
    long long local_resync;  // full solve after this many local updates
    int local_max;           // largest avalanche updated locally
};

// per-replica pointers of the batched stress solve
struct SolveRep {
    const double *eps0;
    void *spec, *sigma;
    double *spart, *res;
    const double *epsavg;
};

struct StzResident;

struct stz_ctx {
    stz_params p;
    StzConsts k;
    int C, S, M, dim;
    int n[3];
    long long N;
    int lc;                 // loaded component
    int prec;               // STZ_FP64 | STZ_MIXED
    size_t rsz;             // sizeof real storage type (8 or 4)
    int h0, hp;             // half spectrum length, padded row stride (complex elements)
    long long nblk;         // rate blocks per mode
    long long ngrp;         // selection groups

    hipStream_t stream = nullptr;

    // device fields
    double *eps0 = nullptr;       // [C][N]
    void *sigma = nullptr;        // [C][N] R
    void *q0 = nullptr;           // [M][N] R
    void *q1 = nullptr;           // [M][N] R, the stress-free barriers a print step restores
                                  // (STZ/init.c:189, STZ/bmg.c:215-223); crystal-like MRO contexts only
    void *deps = nullptr;         // [M][S][N] R
    void *deps_last = nullptr;    // [S][N] R
    double *perm = nullptr, *temp = nullptr, *tlast = nullptr;
    double *tloc = nullptr;       // [N] local temperature (Tlocal), FP64
    uint8_t *mro = nullptr;       // [N] crystal-like MRO type (nullptr: glass everywhere)
    std::vector<uint8_t> mro_h;   // host copy of the map (static after initialisation)
    long long nmro = 0;           // crystal-like voxels
    int heat = 0;                 // HEAT_CONDUCT: diffuse tloc after thermal and elastic steps
    void *spec = nullptr;         // [C][n1*n2][hp] cpx<R>
    void *minq = nullptr;         // [N] R
    void *q0eff = nullptr;        // [M][N] R, cached Q0*exp(-D*S)
    void *scache = nullptr;       // [N] R, softening of the cache entries (NaN: invalid)
    uint8_t *argm = nullptr;      // [N]
    double *bsum = nullptr;       // [M*nblk] block rate sums
    double *dspec = nullptr;      // unused placeholder (kept zero)
    double *hspec = nullptr;      // pinned, device-visible: dice + thermal redraw [1 + M*S]
    double *dspec_dev = nullptr;  // device address of hspec
    double *hath = nullptr;       // pinned, device-visible: athermal redraws [acap_max * M*S]
    double *dath_dev = nullptr;   // device address of hath
    long long acap = 16;          // voxels of an avalanche transformed without a host round trip
    long long acap_max = 256;
    double *gsum = nullptr;       // [ngrp]
    double *csum = nullptr;       // [nchr] chunk totals: NT consecutive group sums each
    int *ccnt = nullptr;          // [ncha] athermal counts per chunk of NT voxel blocks
    long long *gneg = nullptr;    // [ngrp]
    double *res = nullptr;        // small result area (device)
    double *hres = nullptr;       // pinned host mirror
    double *spart = nullptr;      // stress-mean partials [nrowblk][C]
    int *acnt = nullptr;          // athermal per-block counts [nblk]
    int *aoff = nullptr;          // athermal per-block offsets [nblk]
    int *alist = nullptr;         // athermal voxel list [N]
    double *dnew = nullptr;       // new transformation strains for apply [cap]
    size_t dnew_cap = 0;
    double *socp = nullptr;       // athermal DWork partials
    FftTwiddles tw;
    int n_athermal = 0;

    // host-side scalars (the reference's globals)
    double epsavg[6] = {0, 0, 0, 0, 0, 0};
    double stress[6] = {0, 0, 0, 0, 0, 0};
    double totime = 0.0;
    double rate_total = 0.0;
    int pend_comp = -1;
    double pend_delta = 0.0;       // FP64: applied and written back by the next rate pass
    double soff[6] = {0, 0, 0, 0, 0, 0}; // MIXED: uniform stress offset added on read

    // run control: print block and termination tests after every step once stz_set_print
    // has been called (print_on)
    int print_on = 0;
    int printperiod = 0;
    double endstrain = 1e300;
    double strainhist_last = 0.0;
    long long step = 0, astep = 0, kstep = 0, estep = 0, ntrans = 0, nsolve = 0;
    long long avlan = 0;
    double soc_elast = 0.0;

    // event-local stress update (stz_local.hip): impulse response [S][C][N] in the
    // storage precision, shared by contexts of equal shape and elasticity
    int solver_mode = STZ_SOLVER_FFT;
    long long local_resync = 1024;
    int local_max = 16;
    void *kresp = nullptr;
    int *lsrc = nullptr;          // device list of update sites [STZ_LOCAL_MAXSRC]
    double ebase[6] = {0, 0, 0, 0, 0, 0};   // mean strain the field was last solved with
    double usum[6] = {0, 0, 0, 0, 0, 0};    // FP64: uniform increments added to the field since
    long long nlocal = 1ll << 60;           // local updates since the last full solve (huge: none allowed)

    // elastic inhomogeneity (stz_inhom.hip): Lambda(x), Mu(x) in FP64; nullptr = uniform
    double *lamx = nullptr, *mux = nullptr;
    double *tau0 = nullptr;                 // C(x):eps0 [C][N]
    double *estar[2] = {nullptr, nullptr};  // reference-medium eigenstrain, in/out of an iteration [C][N]
    double *cpart = nullptr;                // per-block sums of the squared eigenstrain change [rows]
    int inhom_cur = 0;                      // estar slot solved last (the reference's epsolast)
    int inhom_iters = 1;                    // homogeneous solves of the last stress solve

    int tune[4] = {0, 0, 0, 0};     // solve tile choices (x fwd rows, y width, fused width, x inv rows); 0: default
    long long tune_epoch = 0;       // bumped when the choices change (captured graphs are rebuilt)
    StzRng rng;
    long long rng_epoch = 0;       // bumped when the stream is reseeded or replaced
    StzResident *rs = nullptr;     // resident-step resources (created on first use)
    std::vector<double> hbuf;      // host staging for redraws
    std::string err;

    // kernel timing with HIP events on `stream`
    int timing = 0;
    hipEvent_t tev[24] = {};
    double tms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long tn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double sms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long sn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// event slots: solve 0..6, rates 8..10, select 11..12, apply 13..14, athermal 15..16
inline void tmark(stz_ctx *c, int i)
{
    if (c->timing) (void)hipEventRecord(c->tev[i], c->stream);
}

// ---------------------------------------------------------------------------
// error helpers
// ---------------------------------------------------------------------------
#define STZ_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return STZ_E_HIP;                                                      \
        }                                                                          \
    } while (0)

#define STZ_CHECK_LAUNCH()                                                         \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string("launch: ") + hipGetErrorString(e_);            \
            return STZ_E_HIP;                                                      \
        }                                                                          \
    } while (0)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by s*i (s = +1 or -1)
template <typename R>
__device__ __forceinline__ cpx<R> mul_si(cpx<R> a, R s) { return {-s * a.y, s * a.x}; }

// In-register DFT of size Q, sign s (-1 forward, +1 backward).
template <typename R, int Q>
struct Dft;

template <typename R>
struct Dft<R, 2> {
    static __device__ __forceinline__ void run(cpx<R> *v, R)
    {
        cpx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename R>
struct Dft<R, 4> {
    static __device__ __forceinline__ void run(cpx<R> *v, R s)
    {
        cpx<R> ac = cadd(v[0], v[2]), am = csub(v[0], v[2]);
        cpx<R> bd = cadd(v[1], v[3]), bm = mul_si(csub(v[1], v[3]), s);
        v[0] = cadd(ac, bd);
        v[2] = csub(ac, bd);
        v[1] = cadd(am, bm);
        v[3] = csub(am, bm);
    }
};

template <typename R>
struct Dft<R, 8> {
    static __device__ __forceinline__ void run(cpx<R> *v, R s)
    {
        const R h = (R)0.70710678118654752440084436210484903928;
        cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<R, 4>::run(e, s);
        Dft<R, 4>::run(o, s);
        // o[k] *= w8^k, w8 = e^{s*2*pi*i/8}
        cpx<R> o1 = {h * (o[1].x - s * o[1].y), h * (o[1].y + s * o[1].x)};
        cpx<R> o2 = mul_si(o[2], s);
        cpx<R> o3 = {h * (-o[3].x - s * o[3].y), h * (-o[3].y + s * o[3].x)};
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename R>
struct Dft<R, 3> {
    static __device__ __forceinline__ void run(cpx<R> *v, R s)
    {
        const R c3 = (R)0.86602540378443864676372317075293618347;  // sqrt(3)/2
        cpx<R> a = v[0], b = v[1], c = v[2];
        cpx<R> bc = cadd(b, c);
        cpx<R> t1 = {a.x - (R)0.5 * bc.x, a.y - (R)0.5 * bc.y};
        cpx<R> d = csub(b, c);
        cpx<R> t2 = {-s * c3 * d.y, s * c3 * d.x};
        v[0] = cadd(a, bc);
        v[1] = cadd(t1, t2);
        v[2] = csub(t1, t2);
    }
};

// Radix-16 DFT as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2): DFT-4 over n1 for each n2,
// multiply by w16^(n2 k1), DFT-4 over n2 for each k1; w16 = e^{s*2*pi*i/16}.
template <typename R>
struct Dft<R, 16> {
    // x * w16^m for m = n2*k1 in {1, 2, 3, 4, 6, 9}
    template <int M>
    static __device__ __forceinline__ cpx<R> tw16(cpx<R> x, R s)
    {
        const R c1 = (R)0.92387953251128675612818318939678828682;   // cos(pi/8)
        const R s1 = (R)0.38268343236508977172845998403039886676;   // sin(pi/8)
        const R h = (R)0.70710678118654752440084436210484903928;
        if constexpr (M == 4) {
            return mul_si(x, s);
        } else if constexpr (M == 2 || M == 6) {
            // (+-h, s h) for angle pi/4 and 3 pi/4
            const R c = M == 2 ? h : -h;
            return {c * x.x - s * h * x.y, c * x.y + s * h * x.x};
        } else {
            // cos, sin of 2 pi m/16 for m = 1, 3, 9
            const R c = M == 1 ? c1 : (M == 3 ? s1 : -c1);
            const R sn = M == 1 ? s1 : (M == 3 ? c1 : -s1);
            return {c * x.x - s * sn * x.y, c * x.y + s * sn * x.x};
        }
    }
    static __device__ __forceinline__ void run(cpx<R> *v, R s)
    {
        cpx<R> a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            cpx<R> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            Dft<R, 4>::run(t, s);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
        }
        a[1][1] = tw16<1>(a[1][1], s);
        a[1][2] = tw16<2>(a[1][2], s);
        a[1][3] = tw16<3>(a[1][3], s);
        a[2][1] = tw16<2>(a[2][1], s);
        a[2][2] = tw16<4>(a[2][2], s);
        a[2][3] = tw16<6>(a[2][3], s);
        a[3][1] = tw16<3>(a[3][1], s);
        a[3][2] = tw16<6>(a[3][2], s);
        a[3][3] = tw16<9>(a[3][3], s);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<R> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            Dft<R, 4>::run(t, s);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    }
};

// Radix plans.  A plan code lists the radices of a transform's passes, one per byte,
// first pass in the lowest byte (at most four passes).  Plan 0 of a length is the default
// (radix 8 first, then 4, 2, 3); plans 1..STZ_NPLAN-1 are alternative factorisations the
// autotuner may time (code 0: none for that length): fewer passes (radix 16), or the
// radices in another order.
#define STZ_NPLAN 4
__host__ __device__ constexpr unsigned plan_pack(int a, int b = 0, int c = 0, int d = 0)
{
    return (unsigned)a | ((unsigned)b << 8) | ((unsigned)c << 16) | ((unsigned)d << 24);
}
__host__ __device__ constexpr unsigned default_plan(int L)
{
    unsigned code = 0;
    int sh = 0;
    while (L % 8 == 0 && L > 4) { code |= 8u << sh; sh += 8; L /= 8; }
    while (L % 4 == 0) { code |= 4u << sh; sh += 8; L /= 4; }
    while (L % 2 == 0) { code |= 2u << sh; sh += 8; L /= 2; }
    while (L % 3 == 0) { code |= 3u << sh; sh += 8; L /= 3; }
    return (L == 1 && sh <= 32) ? code : 0u;
}
__host__ __device__ constexpr unsigned plan_code(int L, int j)
{
    if (j == 0) return default_plan(L);
    switch (L) {
    case 16: return j == 1 ? plan_pack(16) : (j == 2 ? plan_pack(4, 4) : 0u);
    case 32: return j == 1 ? plan_pack(16, 2) : (j == 2 ? plan_pack(4, 8) : 0u);
    case 64: return j == 1 ? plan_pack(16, 4) : (j == 2 ? plan_pack(4, 16) : (j == 3 ? plan_pack(8, 4, 2) : 0u));
    case 96: return j == 1 ? plan_pack(4, 8, 3) : (j == 2 ? plan_pack(16, 2, 3) : 0u);
    case 128: return j == 1 ? plan_pack(16, 8) : (j == 2 ? plan_pack(8, 16) : (j == 3 ? plan_pack(8, 4, 4) : 0u));
    case 256: return j == 1 ? plan_pack(16, 16) : (j == 2 ? plan_pack(4, 8, 8) : (j == 3 ? plan_pack(8, 16, 2) : 0u));
    case 512: return j == 1 ? plan_pack(16, 8, 4) : (j == 2 ? plan_pack(8, 16, 4) : (j == 3 ? plan_pack(16, 16, 2) : 0u));
    default: return 0u;
    }
}
__host__ __device__ constexpr int plan_passes(unsigned code)
{
    int c = 0;
    while (c < 4 && ((code >> (8 * c)) & 255u)) ++c;
    return c;
}
__host__ __device__ constexpr int plan_radix(unsigned code, int i) { return (int)((code >> (8 * i)) & 255u); }
__host__ __device__ constexpr long long plan_product(unsigned code)
{
    long long p = 1;
    for (int i = 0; i < plan_passes(code); ++i) p *= plan_radix(code, i);
    return p;
}
// plan j is compiled for length L (plan 0 always, when L is supported)
__host__ __device__ constexpr bool plan_ok(int L, int j)
{
    return j >= 0 && j < STZ_NPLAN && plan_code(L, j) != 0u && plan_product(plan_code(L, j)) == L;
}

// One Stockham pass of radix Q (stride parameter NS = product of the earlier radices)
// over NL lines of length L held in LDS with layout buf[point * LS + line]; in place,
// through registers (two barriers).
template <typename R, int L, int NL, int LS, int NT, int Q, int NS>
__device__ __forceinline__ void fft_pass(cpx<R> *buf, const cpx<R> *tw, R s)
{
    constexpr int NQ = L / Q;
    constexpr int TASKS = NL * NQ;
    constexpr int TPT = (TASKS + NT - 1) / NT;
    constexpr int TSTEP = L / (NS * Q);
    cpx<R> v[TPT][Q];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int t = tid + k * NT;
        if ((TASKS % NT == 0) || t < TASKS) {
            const int line = t % NL, j = t / NL;
#pragma unroll
            for (int r = 0; r < Q; ++r) v[k][r] = buf[(j + r * NQ) * LS + line];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int t = tid + k * NT;
        if ((TASKS % NT == 0) || t < TASKS) {
            const int line = t % NL, j = t / NL;
            const int jm = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < Q; ++r) {
                    cpx<R> w = tw[jm * TSTEP * r];
                    w.y = w.y * (-s);     // table holds e^{-i..}; s = +1 conjugates
                    v[k][r] = cmul(v[k][r], w);
                }
            }
            Dft<R, Q>::run(v[k], s);
            const int idxD = (j / NS) * NS * Q + jm;
#pragma unroll
            for (int r = 0; r < Q; ++r) buf[(idxD + r * NS) * LS + line] = v[k][r];
        }
    }
    __syncthreads();
}

template <typename R, int L, int NL, int LS, int NT, unsigned PC, int I, int NS>
__device__ __forceinline__ void fft_plan(cpx<R> *buf, const cpx<R> *tw, R s)
{
    if constexpr (I < plan_passes(PC)) {
        constexpr int Q = plan_radix(PC, I);
        fft_pass<R, L, NL, LS, NT, Q, NS>(buf, tw, s);
        fft_plan<R, L, NL, LS, NT, PC, I + 1, NS * Q>(buf, tw, s);
    }
}

// Full FFT (natural order in and out) of NL lines of length L, sign s (-1 fwd, +1 bwd),
// by radix plan PL of the length.
template <typename R, int L, int NL, int LS, int NT, int PL = 0>
__device__ __forceinline__ void fft_lines(cpx<R> *buf, const cpx<R> *tw, R s)
{
    static_assert(plan_ok(L, PL), "unsupported FFT length or radix plan");
    fft_plan<R, L, NL, LS, NT, plan_code(L, PL), 0, 1>(buf, tw, s);
}

// fft_pass whose inputs come from ld(line, point) when GIN (the first pass of a transform
// of data in global memory: the points go straight into the registers of the butterflies,
// all loads of a thread issued before the first use) and whose outputs go to
// st(line, point, value) when GOUT (the last pass: no LDS round trip before the store).
// The LDS tile is only touched by the passes in between.
template <typename R, int L, int NL, int LS, int NT, int Q, int NS, bool GIN, bool GOUT, typename Ld, typename St>
__device__ __forceinline__ void fft_pass_io(cpx<R> *buf, const cpx<R> *tw, R s, Ld &ld, St &st)
{
    constexpr int NQ = L / Q;
    constexpr int TASKS = NL * NQ;
    constexpr int TPT = (TASKS + NT - 1) / NT;
    constexpr int TSTEP = L / (NS * Q);
    cpx<R> v[TPT][Q];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int t = tid + k * NT;
        if ((TASKS % NT == 0) || t < TASKS) {
            const int line = t % NL, j = t / NL;
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                if constexpr (GIN) v[k][r] = ld(line, j + r * NQ);
                else v[k][r] = buf[(j + r * NQ) * LS + line];
            }
        }
    }
    if constexpr (!GIN) __syncthreads();   // in place: every read of the tile before a write
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int t = tid + k * NT;
        if ((TASKS % NT == 0) || t < TASKS) {
            const int line = t % NL, j = t / NL;
            const int jm = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < Q; ++r) {
                    cpx<R> w = tw[jm * TSTEP * r];
                    w.y = w.y * (-s);
                    v[k][r] = cmul(v[k][r], w);
                }
            }
            Dft<R, Q>::run(v[k], s);
            const int idxD = (j / NS) * NS * Q + jm;
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                if constexpr (GOUT) st(line, idxD + r * NS, v[k][r]);
                else buf[(idxD + r * NS) * LS + line] = v[k][r];
            }
        }
    }
    if constexpr (!GOUT) __syncthreads();
}

// passes I..IEND-1 of plan PC (the stride parameter NS of pass I given); GIN applies to
// pass 0 and GOUT to the plan's last pass
template <typename R, int L, int NL, int LS, int NT, unsigned PC, int I, int IEND, int NS, bool GIN, bool GOUT,
          typename Ld, typename St>
__device__ __forceinline__ void fft_plan_range(cpx<R> *buf, const cpx<R> *tw, R s, Ld &ld, St &st)
{
    if constexpr (I < IEND) {
        constexpr int Q = plan_radix(PC, I);
        constexpr bool LASTP = I + 1 == plan_passes(PC);
        fft_pass_io<R, L, NL, LS, NT, Q, NS, GIN && I == 0, GOUT && LASTP>(buf, tw, s, ld, st);
        fft_plan_range<R, L, NL, LS, NT, PC, I + 1, IEND, NS * Q, GIN, GOUT>(buf, tw, s, ld, st);
    }
}

template <typename R, int L, int NL, int LS, int NT, unsigned PC, int I, int NS, bool GIN, bool GOUT, typename Ld,
          typename St>
__device__ __forceinline__ void fft_plan_io(cpx<R> *buf, const cpx<R> *tw, R s, Ld &ld, St &st)
{
    fft_plan_range<R, L, NL, LS, NT, PC, I, plan_passes(PC), NS, GIN, GOUT>(buf, tw, s, ld, st);
}

// fft_lines with the input read through ld (GIN) and/or the output written through st
// (GOUT) instead of the LDS tile; the twiddle table must be in LDS before the second pass
// (the first pass multiplies by no twiddle)
template <typename R, int L, int NL, int LS, int NT, bool GIN, bool GOUT, int PL = 0, typename Ld, typename St>
__device__ __forceinline__ void fft_lines_io(cpx<R> *buf, const cpx<R> *tw, R s, Ld ld, St st)
{
    static_assert(plan_ok(L, PL), "unsupported FFT length or radix plan");
    fft_plan_io<R, L, NL, LS, NT, plan_code(L, PL), 0, 1, GIN, GOUT>(buf, tw, s, ld, st);
}

// <sigma> from the stress-solve block partials spart[b][c] (fixed order: per-thread
// strided runs, then a tree); a workgroup of exactly STZ_NT_MEAN threads; out[c] for c < C
#define STZ_NT_MEAN 1024
template <int C>
__device__ void mean_partials(const double *spart, int nblocks, double invN, double *out)
{
    __shared__ double red[C][STZ_NT_MEAN];
    const int tid = threadIdx.x;
    double a[C];
#pragma unroll
    for (int c = 0; c < C; ++c) a[c] = 0.0;
    for (int b = tid; b < nblocks; b += STZ_NT_MEAN)
#pragma unroll
        for (int c = 0; c < C; ++c) a[c] += spart[(long long)b * C + c];
#pragma unroll
    for (int c = 0; c < C; ++c) red[c][tid] = a[c];
    __syncthreads();
    for (int w = STZ_NT_MEAN / 2; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int c = 0; c < C; ++c) red[c][tid] = red[c][tid] + red[c][tid + w];
        __syncthreads();
    }
    if (tid < C) out[tid] = red[tid][0] * invN;
    __syncthreads();
}

// mean_partials on a workgroup of NTH threads, with the very operations of the
// STZ_NT_MEAN-thread tree: thread t keeps the strided runs of the virtual threads
// t + j*NTH, the tree levels whose two operands belong to one thread add in registers
// (v and v + w with w >= NTH), the remaining levels in LDS.  Same value bit for bit.
template <int C, int NTH>
__device__ void mean_partials_n(const double *spart, int nblocks, double invN, double *out)
{
    static_assert(NTH <= STZ_NT_MEAN && STZ_NT_MEAN % NTH == 0, "NTH divides the tree width");
    constexpr int J = STZ_NT_MEAN / NTH;
    __shared__ double red[C][NTH];
    const int tid = threadIdx.x;
    double a[J][C];
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int c = 0; c < C; ++c) a[j][c] = 0.0;
#pragma unroll
    for (int j = 0; j < J; ++j)
        for (int b = tid + j * NTH; b < nblocks; b += STZ_NT_MEAN)
#pragma unroll
            for (int c = 0; c < C; ++c) a[j][c] += spart[(long long)b * C + c];
#pragma unroll
    for (int w = STZ_NT_MEAN / 2; w >= NTH; w >>= 1)
#pragma unroll
        for (int j = 0; j < w / NTH; ++j)
#pragma unroll
            for (int c = 0; c < C; ++c) a[j][c] = a[j][c] + a[j + w / NTH][c];
#pragma unroll
    for (int c = 0; c < C; ++c) red[c][tid] = a[0][c];
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int c = 0; c < C; ++c) red[c][tid] = red[c][tid] + red[c][tid + w];
        __syncthreads();
    }
    if (tid < C) out[tid] = red[tid][0] * invN;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// host-side launchers implemented in the .hip files
// ---------------------------------------------------------------------------
int stz_fft_setup(stz_ctx *ctx);
void stz_fft_free(stz_ctx *ctx);
int stz_solve_launch(stz_ctx *ctx, int gated); // enqueues the full stress solve on ctx->stream
int stz_displacement_launch(stz_ctx *ctx, double *u);   // print-step displacement into u[dim][N] (device)
int stz_solve_spectral(stz_ctx *ctx);   // y passes + fused Gamma-hat pass of ctx->spec (ungated)
// elastic inhomogeneity (stz_inhom.hip)
int stz_inhom_fwd_x(stz_ctx *ctx, const double *src);   // R2C x pass of an eigenstrain [C][N], no round trip
int stz_inhom_run(stz_ctx *ctx);        // the fixed-point loop; <sigma> in res[0..C), sigma written
int stz_inhom_free(stz_ctx *ctx);
// crystal-like MRO map (stz_api.hip): ctx->mro_h to the device, nmro, avalanche drawing
int stz_set_mro_map(stz_ctx *ctx);
// heat conduction (stz_heat.hip): one HeatConducting(hdt) of ctx->tloc, NaN flag in res[15]
int stz_heat_launch(stz_ctx *ctx, double hdt);
int stz_rates_launch(stz_ctx *ctx, int decide); // enqueues rates + group sums (+ step-kind decision)
int stz_select_launch(stz_ctx *ctx, double target, int gated);
int stz_apply_launch(stz_ctx *ctx, int site, int mode, int gated);
int stz_athermal_launch(stz_ctx *ctx);
int stz_athermal_apply_launch(stz_ctx *ctx, int n);
int stz_athermal_fused_launch(stz_ctx *ctx, int gated);   // + the |DWork| sum
int stz_athermal_apply_range(stz_ctx *ctx, long long first, long long count);
int stz_energy_launch(stz_ctx *ctx);
bool stz_solve_supported(int dim, const int *n, int prec);
int tune_candidates(const stz_ctx *ctx, int slot, int out[STZ_TUNE_MAXCAND]);
int stz_init_device_launch(stz_ctx *ctx, unsigned long long seed, const double *spheres_host, int nsph, double inc_mean,
                           double inc_std, long long *inside);
// batched solve without the mean kernel (the caller reduces the *nparts partials)
int stz_solve_launch_batched(stz_ctx *proto, int nrep, const SolveRep *reps, hipStream_t stream, int *nparts);

// marks the derived per-voxel arrays stale after a restore (the next rate pass rebuilds them)
int stz_invalidate_derived(stz_ctx *ctx);

// event-local stress update (stz_local.hip)
int stz_local_prepare(stz_ctx *ctx);    // builds (or takes from the cache) the folded impulse response
void stz_local_release(stz_ctx *ctx);
void stz_local_mark_solved(stz_ctx *ctx); // a full solve rewrote the field with ctx->epsavg
void stz_local_uniform(const stz_ctx *ctx, double *uni);
int stz_local_launch(stz_ctx *ctx, int ns, const double *uni);   // sites in ctx->lsrc
long long stz_local_blocks(const stz_ctx *ctx);

// resident step (stz_resident.hip)
struct StzPlan;
void stz_rng_sync(stz_ctx *ctx);        // moves ctx->rng to the resident stream position
void stz_rng_replaced(stz_ctx *ctx);    // the host stream was reseeded or overwritten
void stz_resident_free(stz_ctx *ctx);
int stz_plan_create(stz_ctx *const *ctxs, int n, StzPlan **out);
void stz_plan_destroy(StzPlan *pl);
void stz_plan_config(StzPlan *pl, int batch, int use_graph);
int stz_plan_run(StzPlan *pl, int nsteps, stz_step_result *const *results, int *done, int *rcs, int stop_at_print);
StzPlan *stz_self_plan(stz_ctx *ctx);
