// stz_rng.h -- the reference's random stream: GSL "mt19937" (gsl_rng_default, seeded as
// STZ/init.c:142-150) and gsl_ran_gaussian's polar method (restated at STZ/init.c:207-223).
#pragma once
#include <cstdint>
#include <vector>

struct StzRng {
    uint32_t state[624];
    int pos = 624;
    long long count = 0;      // words produced since seeding (stream position)

    void seed(unsigned long s);
    uint32_t next_u32();
    void skip(long long n);   // advance by n words
    // gsl_rng_uniform: u32 / 2^32, in [0, 1)
    double uniform() { return next_u32() / 4294967296.0; }
    // gsl_rng_uniform_pos: (0, 1)
    double uniform_pos();
    // gsl_ran_gaussian(r, sigma)
    double gaussian(double sigma);
};

// The polar method's value for an accepted pair with second coordinate b and radius^2 rr:
// sigma * b * sqrt(-2 ln rr / rr), evaluated left to right as the reference.
double stz_polar_value(double sigma, double b, double rr);

// Lookup tables over a window [base, base + len) of the stream, so that a kernel can
// take any random number of the reference's sequence by index instead of drawing it.
//
// A gaussian draw consumes word pairs (x, y) until x^2 + y^2 lies in (0, 1] (with
// x = -1 + 2u, u = word/2^32).  Starting from a word of parity p (relative to base),
// consecutive attempts use the pairs (p + 2j, p + 2j + 1), j = 0, 1, ...: the draws of
// such a start are the accepted pairs of parity p in order.  So for p = 0, 1 the
// window keeps the accepted values g[p][k], the pair index pair[p][k] of the k-th
// accepted pair, and rank[p][j] = number of accepted pairs before pair j.  A start
// at relative word r takes p = r & 1, k0 = rank[p][r >> 1]; its c draws are
// g[p][k0 .. k0+c) and the stream continues at p + 2 (pair[p][k0+c-1] + 1).
// gsl_rng_uniform_pos re-draws a zero word, which breaks the pairing: the window ends
// before the first zero word (probability 2^-32 per word).
struct StzMtWindow {
    StzRng at_base;                     // generator state at `base`
    long long base = 0;
    long long len = 0;                  // usable words
    std::vector<uint32_t> words;
    std::vector<double> g[2];
    std::vector<int> rank[2];           // [(len - p) / 2 + 1]
    std::vector<int> pair[2];
    long long nacc[2] = {0, 0};

    // builds the window of T words that starts at the current position of `start`
    void build(const StzRng &start, long long T, int threads);
};
