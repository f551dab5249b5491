// stz_philox.h -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11) for the device-resident paths: the
// ensemble and the initial state of runs that do not replay the reference's MT19937
// stream.  Every draw is a pure function of (key, counter), so any voxel, mode or step
// can generate its numbers independently and in any order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct Philox4 {
    uint32_t v[4];
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

// 10 rounds; the key is bumped by the Weyl constants between rounds
__host__ __device__ inline Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(M0, c.v[0], hi0, lo0);
        philox_mulhilo(M1, c.v[2], hi1, lo1);
        Philox4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
    }
    return c;
}

// 53-bit uniform in [0, 1) from two words
__host__ __device__ inline double philox_u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// two independent N(0,1) draws (Box-Muller on two 53-bit uniforms) for one counter
__device__ inline void philox_normal2(uint32_t key0, uint32_t key1, uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3, double &g0, double &g1)
{
    const Philox4 r = philox4x32_10(Philox4{{c0, c1, c2, c3}}, key0, key1);
    const double u1 = 1.0 - philox_u53(r.v[0], r.v[1]);   // (0, 1]
    const double u2 = philox_u53(r.v[2], r.v[3]);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    g0 = rad * c;
    g1 = rad * s;
}

// purposes of the third counter word
#define STZ_PHX_Q0 0x51300000u        // initial barrier per (mode, voxel)
#define STZ_PHX_DEPS 0x44450000u      // initial transformation strain per (mode, voxel, pair)
#define STZ_PHX_DICE 0x44494345u      // selection uniform per step
#define STZ_PHX_REDRAW 0x52440000u    // thermal-event redraw per (step, mode, pair)
#define STZ_PHX_ATH 0x41540000u       // athermal redraw per (step, voxel, mode, pair)
