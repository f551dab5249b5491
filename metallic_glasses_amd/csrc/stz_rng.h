// stz_rng.h -- the reference's random stream: GSL "mt19937" (gsl_rng_default, seeded as
// STZ/init.c:142-150) and gsl_ran_gaussian's polar method (restated at STZ/init.c:207-223).
#pragma once
#include <cstdint>

struct StzRng {
    uint32_t state[624];
    int pos = 624;

    void seed(unsigned long s);
    uint32_t next_u32();
    // gsl_rng_uniform: u32 / 2^32, in [0, 1)
    double uniform() { return next_u32() / 4294967296.0; }
    // gsl_rng_uniform_pos: (0, 1)
    double uniform_pos();
    // gsl_ran_gaussian(r, sigma)
    double gaussian(double sigma);
};
