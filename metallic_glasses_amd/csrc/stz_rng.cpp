// stz_rng.cpp -- host MT19937 (init_genrand seeding, GSL's seed 0 -> 4357) and the
// polar Gaussian.  The parity path draws every random number here, in the order the
// reference does (STZ/init.c:355-396, 181-198; STZ/elastic.c:124-129, 508-512;
// STZ/evolution.c:142).
#include "stz_rng.h"

#include <cmath>

void StzRng::seed(unsigned long s)
{
    if (s == 0) s = 4357;
    state[0] = static_cast<uint32_t>(s);
    for (int i = 1; i < 624; ++i) {
        uint32_t prev = state[i - 1];
        state[i] = 1812433253u * (prev ^ (prev >> 30)) + static_cast<uint32_t>(i);
    }
    pos = 624;
}

uint32_t StzRng::next_u32()
{
    if (pos >= 624) {
        // regenerate the whole block: twist with the word 397 ahead (cyclically)
        for (int i = 0; i < 624; ++i) {
            uint32_t hi = state[i] & 0x80000000u;
            uint32_t lo = state[(i + 1) % 624] & 0x7fffffffu;
            uint32_t y = hi | lo;
            uint32_t x = state[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) x ^= 0x9908b0dfu;
            state[i] = x;
        }
        pos = 0;
    }
    uint32_t z = state[pos++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
}

double StzRng::uniform_pos()
{
    double u;
    do {
        u = uniform();
    } while (u == 0.0);
    return u;
}

double StzRng::gaussian(double sigma)
{
    double a, b, rr;
    for (;;) {
        a = -1 + 2 * uniform_pos();
        b = -1 + 2 * uniform_pos();
        rr = a * a + b * b;
        if (rr <= 1.0 && rr != 0) break;
    }
    // sigma * y * sqrt(-2 ln r^2 / r^2), evaluated left to right as the reference
    return sigma * b * std::sqrt(-2.0 * std::log(rr) / rr);
}
