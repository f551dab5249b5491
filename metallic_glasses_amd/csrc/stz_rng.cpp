// stz_rng.cpp -- host MT19937 (init_genrand seeding, GSL's seed 0 -> 4357) and the
// polar Gaussian.  The parity path draws every random number here, in the order the
// reference does (STZ/init.c:355-396, 181-198; STZ/elastic.c:124-129, 508-512;
// STZ/evolution.c:142).
#include "stz_rng.h"

#include <algorithm>
#include <cmath>
#include <thread>

void StzRng::seed(unsigned long s)
{
    if (s == 0) s = 4357;
    state[0] = static_cast<uint32_t>(s);
    for (int i = 1; i < 624; ++i) {
        uint32_t prev = state[i - 1];
        state[i] = 1812433253u * (prev ^ (prev >> 30)) + static_cast<uint32_t>(i);
    }
    pos = 624;
    count = 0;
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
    ++count;
    uint32_t z = state[pos++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
}

void StzRng::skip(long long n)
{
    for (long long i = 0; i < n; ++i) (void)next_u32();
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
    return stz_polar_value(sigma, b, rr);
}

double stz_polar_value(double sigma, double b, double rr)
{
    // sigma * y * sqrt(-2 ln r^2 / r^2), evaluated left to right as the reference
    return sigma * b * std::sqrt(-2.0 * std::log(rr) / rr);
}

namespace {

// accept test of one attempt, with the arithmetic of StzRng::gaussian
inline bool polar_accept(uint32_t wx, uint32_t wy, double &b, double &rr)
{
    const double a = -1 + 2 * (wx / 4294967296.0);
    b = -1 + 2 * (wy / 4294967296.0);
    rr = a * a + b * b;
    return rr <= 1.0 && rr != 0;
}

}  // namespace

void StzMtWindow::build(const StzRng &start, long long T, int threads)
{
    at_base = start;
    base = start.count;
    StzRng gen = start;
    words.resize((size_t)T);
    len = T;
    for (long long i = 0; i < T; ++i) {
        words[(size_t)i] = gen.next_u32();
        if (words[(size_t)i] == 0 && len == T) len = i;
    }
    if (threads < 1) threads = 1;
    for (int p = 0; p < 2; ++p) {
        const long long np = len > p ? (len - p) / 2 : 0;
        rank[p].assign((size_t)np + 1, 0);
        // pass 1: accept flags and per-chunk counts, in parallel
        std::vector<uint8_t> acc((size_t)np);
        const long long chunk = (np + threads - 1) / threads;
        std::vector<long long> cnt((size_t)threads, 0);
        auto flags = [&](int t) {
            const long long j0 = (long long)t * chunk, j1 = std::min(np, j0 + chunk);
            long long c = 0;
            for (long long j = j0; j < j1; ++j) {
                double b, rr;
                acc[(size_t)j] = polar_accept(words[(size_t)(p + 2 * j)], words[(size_t)(p + 2 * j + 1)], b, rr);
                c += acc[(size_t)j];
            }
            cnt[(size_t)t] = c;
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(flags, t);
        flags(0);
        for (auto &th : pool) th.join();
        pool.clear();
        std::vector<long long> first((size_t)threads + 1, 0);
        for (int t = 0; t < threads; ++t) first[(size_t)t + 1] = first[(size_t)t] + cnt[(size_t)t];
        nacc[p] = first[(size_t)threads];
        g[p].resize((size_t)nacc[p]);
        pair[p].resize((size_t)nacc[p]);
        // pass 2: ranks, pair indices and values
        auto fill = [&](int t) {
            const long long j0 = (long long)t * chunk, j1 = std::min(np, j0 + chunk);
            long long k = first[(size_t)t];
            for (long long j = j0; j < j1; ++j) {
                rank[p][(size_t)j] = (int)k;
                if (acc[(size_t)j]) {
                    double b, rr;
                    (void)polar_accept(words[(size_t)(p + 2 * j)], words[(size_t)(p + 2 * j + 1)], b, rr);
                    g[p][(size_t)k] = stz_polar_value(1.0, b, rr);
                    pair[p][(size_t)k] = (int)j;
                    ++k;
                }
            }
        };
        for (int t = 1; t < threads; ++t) pool.emplace_back(fill, t);
        fill(0);
        for (auto &th : pool) th.join();
        rank[p][(size_t)np] = (int)nacc[p];
    }
}
