// rng_check.cpp -- the library's host random stream (metallic_glasses_amd/csrc/stz_rng.cpp)
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_oracle_sanitized.py compiles and runs it on the CPU).  Checks: the MT19937
// words equal std::mt19937's from the same seed; skip(n) equals n draws; for starts spread
// over a lookup window built on three threads, the window's Gaussian values equal the
// generator's own draws from that position, and the recorded pair indices give the
// stream position after them.  Exit status 0 and "ok" on success.
#include <cstdio>
#include <random>

#include "stz_rng.h"

static int fail(const char *what, long long a, long long b)
{
    std::fprintf(stderr, "rng_check: %s (%lld, %lld)\n", what, a, b);
    return 1;
}

int main()
{
    StzRng r;
    r.seed(5489);
    std::mt19937 ref(5489);
    for (int i = 0; i < 20000; ++i) {
        const uint32_t x = r.next_u32(), y = ref();
        if (x != y) return fail("mt19937 word", i, (long long)x - (long long)y);
    }

    StzRng a, b;
    a.seed(1234);
    b.seed(1234);
    a.skip(12345);
    for (int i = 0; i < 12345; ++i) b.next_u32();
    for (int i = 0; i < 100; ++i)
        if (a.next_u32() != b.next_u32()) return fail("skip", i, 0);

    StzRng s;
    s.seed(777);
    s.skip(1001);
    StzMtWindow w;
    const long long T = 200000;
    w.build(s, T, 3);
    if (w.base != 1001 || w.len < T - 2) return fail("window extent", w.base, w.len);
    long long checked = 0;
    for (long long start = 0; start + 2000 < w.len; start += 997) {
        for (int shift = 0; shift < 2; ++shift) {
            const long long rel = start + shift;
            const int p = (int)(rel & 1);
            const long long k0 = w.rank[p][(size_t)(rel >> 1)];
            StzRng q = w.at_base;
            q.skip(rel);
            for (int c = 0; c < 40; ++c) {
                if (k0 + c >= w.nacc[p]) return fail("window too short", rel, c);
                const double x = q.gaussian(1.0);
                if (x != w.g[p][(size_t)(k0 + c)]) return fail("gaussian value", rel, c);
                const long long next = p + 2 * ((long long)w.pair[p][(size_t)(k0 + c)] + 1);
                if (q.count - w.base != next) return fail("stream position", rel, c);
                ++checked;
            }
        }
    }
    std::printf("ok: %lld window draws\n", checked);
    return 0;
}
