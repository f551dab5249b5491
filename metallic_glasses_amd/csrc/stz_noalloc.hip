// stz_noalloc.hip -- stz_rates without a context and without any allocation.
//
// The same rate pass as stz_rates (SofteningFactor, ActEnergy, the Q<0 count and the
// KMC rate total, STZ/init.c:711-739, 561-690; STZ/bmg.c:147-157; STZ/evolution.c:120-134)
// over fields the caller owns in device memory, for callers that cannot allocate (code
// under HIP graph capture, tight loops that own their buffers).  Its block and group rate
// sums live in one file-scope device array; each device holds its own copy, whose address
// is looked up once per device with hipGetSymbolAddress and used for every launch.  The
// results go to caller-provided device memory, so the call neither allocates nor
// synchronises.
#include "stz_kmc_dev.h"

#include <cmath>
#include <mutex>
#include <unordered_map>

namespace {

// block rate sums: modes * ceil(voxels / 256) of them at most (16 MiB); group sums of 256
// block sums each; one Q<0 counter (kept in its own 256-byte slot)
constexpr long long kMaxBlockSums = 1ll << 21;
constexpr long long kMaxGroupSums = kMaxBlockSums / STZ_GRP;
constexpr long long kScratchWords = kMaxBlockSums + kMaxGroupSums + 32;

__device__ __align__(256) double scratch[kScratchWords];

std::mutex g_mu;
std::unordered_map<int, double *> g_scratch;   // device ordinal -> address of `scratch` there

// the address of `scratch` on the current device (looked up once per device)
int scratch_on_device(double **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return STZ_E_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) {
        *out = it->second;
        return STZ_OK;
    }
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return STZ_E_HIP;
    g_scratch[dev] = (double *)p;
    *out = (double *)p;
    return STZ_OK;
}

__global__ void k_zero_counter(unsigned long long *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_rates_lean(RateArgs<double> a)
{
    rates_block<double, DIM, false, true>(a, blockIdx.x);
}

__global__ __launch_bounds__(NT) void k_group_lean(const double *__restrict__ bsum, long long nb, double *__restrict__ gsum,
                                                   long long ng)
{
    __shared__ double ws[NW];
    group_sums(bsum, nb, gsum, ng, (long long)blockIdx.x * STZ_GRP_PER_BLOCK, ws);
}

// R by the same running scan over the group sums as the context path (bit-identical)
__global__ __launch_bounds__(NT) void k_total_lean(const double *__restrict__ gsum, long long ng,
                                                   const unsigned long long *__restrict__ nn, double *total_rate,
                                                   long long *n_negative)
{
    __shared__ double ws[NW];
    const double R = group_total(gsum, ng, ws);
    if (threadIdx.x == 0) {
        *total_rate = R;
        *n_negative = (long long)*nn;
    }
}

}  // namespace

extern "C" {

int stz_rates_noalloc(const stz_params *p, const stz_rate_fields *f, double totime, void *stream, double *total_rate,
                      long long *n_negative)
{
    if (!p || !f || !total_rate || !n_negative) return STZ_E_ARG;
    if (p->dim != 2 && p->dim != 3) return STZ_E_ARG;
    if (p->precision != STZ_FP64) return STZ_E_ARG;                    // FP64 fields only
    if (p->modes < 1 || p->modes > STZ_MAXM) return STZ_E_ARG;
    if (!f->sigma || !f->q0 || !f->deps || !f->deps_last || !f->perm || !f->temp || !f->tlast) return STZ_E_ARG;
    const long long N = (long long)p->n[0] * p->n[1] * (p->dim == 2 ? 1 : p->n[2]);
    if (N <= 0) return STZ_E_ARG;
    const long long nblk = (N + STZ_BLK - 1) / STZ_BLK;
    const long long nb = (long long)p->modes * nblk;
    if (nb > kMaxBlockSums) return STZ_E_TOO_LARGE;                    // modes * voxel blocks above the scratch
    const long long ngrp = (nb + STZ_GRP - 1) / STZ_GRP;
    double *sc = nullptr;
    int rc = scratch_on_device(&sc);
    if (rc) return rc;
    double *bsum = sc;
    double *gsum = sc + kMaxBlockSums;
    unsigned long long *nn = reinterpret_cast<unsigned long long *>(sc + kMaxBlockSums + kMaxGroupSums);
    hipStream_t s = (hipStream_t)stream;

    // constants as Init derives them (STZ/init.c:6-75)
    const double kT = p->temperature * STZ_BOLZ * STZ_J_IN_EV;
    const double voxel = p->lx * p->ly * p->lz;
    const double ratepre = voxel * p->trial_frequency;
    RateArgs<double> a;
    a.sigma = const_cast<double *>(f->sigma);     // read only: no pending increment (pend_comp = -1)
    a.q0 = f->q0;
    a.deps = f->deps;
    a.deps_last = f->deps_last;
    a.q0eff = nullptr;
    a.scache = nullptr;
    a.perm = f->perm;
    a.temp = f->temp;
    a.tlast = f->tlast;
    a.minq = nullptr;
    a.argm = nullptr;
    a.mro = nullptr;
    a.bsum = bsum;
    a.acnt = nullptr;
    a.nneg = nn;
    a.thr = p->athermal_threshold;
    a.N = N;
    a.nblk = nblk;
    a.M = p->modes;
    a.softening = p->softening == 1;
    a.pend_comp = -1;
    a.pend_delta = 0.0;
    for (int c = 0; c < 6; ++c) a.soff[c] = 0.0;
    a.totime = totime;
    a.tTemp = 1 / (p->trial_frequency * std::exp(-p->temp_soft_barrier / kT));
    a.kT = kT;
    a.ratepre = ratepre;
    a.voxel = voxel;
    a.lnpre = std::log(ratepre);

    // the counter starts at zero: a fill kernel rather than a memset, so that the sequence
    // can be captured in a graph
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, nn);
    if (p->dim == 2) hipLaunchKernelGGL(k_rates_lean<2>, dim3((unsigned)nblk), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL(k_rates_lean<3>, dim3((unsigned)nblk), dim3(NT), 0, s, a);
    const unsigned gb = (unsigned)((ngrp + STZ_GRP_PER_BLOCK - 1) / STZ_GRP_PER_BLOCK);
    hipLaunchKernelGGL(k_group_lean, dim3(gb), dim3(NT), 0, s, (const double *)bsum, nb, gsum, ngrp);
    hipLaunchKernelGGL(k_total_lean, dim3(1), dim3(NT), 0, s, (const double *)gsum, ngrp,
                       (const unsigned long long *)nn, total_rate, n_negative);
    return hipGetLastError() == hipSuccess ? STZ_OK : STZ_E_HIP;
}

}  // extern "C"
