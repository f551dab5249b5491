// stz_init.hip -- device-generated initial state for large grids.
//
// The reference draws its initial barriers and transformation strains from one MT19937
// stream on the host (Energy_Init, STZ/init.c:181-198; DStrain, STZ/elastic.c:124-129).
// That sequential stream takes minutes at 10^7-10^8 voxels, so runs that do not need the
// reference's exact initial numbers (the mixed-vs-FP64 comparisons at 256^3 and
// 512x512x256) draw the same distributions on the GPU with a counter-based generator
// (Philox4x32-10, stz_philox.h): every value is a function of (seed, voxel, mode,
// component), identical in both precisions and for any launch shape.
//   Q0[m][v]        = mean + std * N(0,1), or inc_mean + inc_std * N(0,1) inside a soft
//                     inclusion (a voxel within radius r of one of the given sphere
//                     centres, periodic distance)
//   deps[m][k][v]   = 0.5 * gamma0 * N(0,1)
// All voxels are glass (the MRO fraction of the configurations is 0).
#include "stz_internal.h"
#include "stz_philox.h"

namespace {

constexpr int NT_INIT = 256;

struct InitArgs {
    void *q0, *deps;
    long long N;
    int n0, n1, n2, M, S;
    uint32_t k0, k1;
    double mean, sd, inc_mean, inc_sd, half_gamma0;
    const double *spheres;   // [nsph][4]: centre (x, y, z) in voxels, radius
    int nsph;
    unsigned long long *inside;   // count of inclusion voxels
};

__device__ __forceinline__ double pdist2(double a, double b, int n)
{
    double d = fabs(a - b);
    if (d > 0.5 * n) d = n - d;
    return d * d;
}

template <typename R>
__global__ __launch_bounds__(NT_INIT) void k_init(InitArgs a)
{
    const long long v = (long long)blockIdx.x * NT_INIT + threadIdx.x;
    if (v >= a.N) return;
    const int i0 = (int)(v % a.n0);
    const int i1 = (int)((v / a.n0) % a.n1);
    const int i2 = (int)(v / ((long long)a.n0 * a.n1));
    bool inc = false;
    for (int s = 0; s < a.nsph && !inc; ++s) {
        const double *c = a.spheres + 4 * s;
        const double d2 = pdist2(i0, c[0], a.n0) + pdist2(i1, c[1], a.n1) + pdist2(i2, c[2], a.n2);
        inc = d2 <= c[3] * c[3];
    }
    if (inc) atomicAdd(a.inside, 1ull);
    const double mean = inc ? a.inc_mean : a.mean, sd = inc ? a.inc_sd : a.sd;
    const uint32_t vl = (uint32_t)v, vh = (uint32_t)(v >> 32);
    R *q0 = (R *)a.q0;
    R *deps = (R *)a.deps;
    for (int m = 0; m < a.M; ++m) {
        double g0, g1;
        philox_normal2(a.k0, a.k1, vl, vh, STZ_PHX_Q0 | (uint32_t)m, 0u, g0, g1);
        q0[(long long)m * a.N + v] = (R)(mean + sd * g0);
        for (int k = 0; k < a.S; k += 2) {
            philox_normal2(a.k0, a.k1, vl, vh, STZ_PHX_DEPS | (uint32_t)m, (uint32_t)k, g0, g1);
            deps[((long long)m * a.S + k) * a.N + v] = (R)(a.half_gamma0 * g0);
            if (k + 1 < a.S) deps[((long long)m * a.S + k + 1) * a.N + v] = (R)(a.half_gamma0 * g1);
        }
    }
}

}  // namespace

int stz_init_device_launch(stz_ctx *ctx, unsigned long long seed, const double *spheres_host, int nsph, double inc_mean,
                           double inc_std, long long *inside)
{
    InitArgs a;
    a.q0 = ctx->q0;
    a.deps = ctx->deps;
    a.N = ctx->N;
    a.n0 = ctx->n[0];
    a.n1 = ctx->n[1];
    a.n2 = ctx->n[2];
    a.M = ctx->M;
    a.S = ctx->S;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.mean = ctx->p.q0_mean;
    a.sd = ctx->p.q0_std;
    a.inc_mean = inc_mean;
    a.inc_sd = inc_std;
    a.half_gamma0 = ctx->k.half_gamma0;
    a.nsph = nsph;
    double *dsph = nullptr;
    unsigned long long *dcnt = nullptr;
    STZ_HIP(hipMalloc(&dcnt, sizeof(unsigned long long)));
    STZ_HIP(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), ctx->stream));
    if (nsph > 0) {
        STZ_HIP(hipMalloc(&dsph, sizeof(double) * 4 * (size_t)nsph));
        STZ_HIP(hipMemcpyAsync(dsph, spheres_host, sizeof(double) * 4 * (size_t)nsph, hipMemcpyHostToDevice, ctx->stream));
    }
    a.spheres = dsph;
    a.inside = dcnt;
    const dim3 grid((unsigned)((ctx->N + NT_INIT - 1) / NT_INIT));
    if (ctx->prec == STZ_FP64) hipLaunchKernelGGL((k_init<double>), grid, dim3(NT_INIT), 0, ctx->stream, a);
    else hipLaunchKernelGGL((k_init<float>), grid, dim3(NT_INIT), 0, ctx->stream, a);
    STZ_CHECK_LAUNCH();
    unsigned long long cnt = 0;
    STZ_HIP(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    STZ_HIP(hipStreamSynchronize(ctx->stream));
    if (dsph) (void)hipFree(dsph);
    (void)hipFree(dcnt);
    if (inside) *inside = (long long)cnt;
    return STZ_OK;
}

extern "C" {

// Host evaluation of the generator's integer rounds (the code the kernels run), for
// known-answer checks of Philox4x32-10.
int stz_philox_probe(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4])
{
    if (!ctr || !key || !out) return STZ_E_ARG;
    const Philox4 r = philox4x32_10(Philox4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
    return STZ_OK;
}

}  // extern "C"
