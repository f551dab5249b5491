/*
 * oracle_check.c -- drives every entry point of the CPU oracle (oracle/stz_oracle.c) on
 * small grids, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_oracle_sanitized.py compiles and runs it on the CPU).  Besides running
 * clean under the sanitizers it checks that a state image written mid-run and loaded into
 * a fresh state continues with the same step records.  Exit status 0 and "ok" on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "stz_oracle.h"

static ora_params reference_2d(void)
{
    /* the reference input's physics (STZ/input, STZ/bmg.h, STZ/init.c:60-68) */
    ora_params p;
    memset(&p, 0, sizeof p);
    p.dim = 2;
    p.n[0] = 32; p.n[1] = 32; p.n[2] = 1;
    p.modes = 20;
    p.softening = 1;
    p.seed = 123;
    p.E = 69.46; p.nu = 0.365; p.double_nu = 1;
    p.q0_mean = 2.0; p.q0_std = 0.25;
    p.gamma0 = 0.14;
    p.temperature = 300.0;
    p.trial_frequency = 1e13;
    p.lx = 0.7; p.ly = 0.7; p.lz = 1.0;
    p.strainrate = 1e-4; p.strainincrement = 1e-4;
    p.perm_soft = 13.0; p.temp_soft = 39.0; p.perm_cap = -log(0.6); p.temp_soft_barrier = 0.37;
    p.athermal_threshold = 0.33;
    p.mro_q0 = 2.0; p.random_ratio = 0.2;
    return p;
}

static int fail(const char *what)
{
    fprintf(stderr, "oracle_check: %s\n", what);
    return 1;
}

/* n steps; counts of thermal and athermal steps; nonzero return on an oracle error */
static int run(ora_state *s, int n, int *thermal, int *athermal, ora_step_info *last)
{
    for (int i = 0; i < n; ++i) {
        ora_step_info info;
        const int rc = ora_step(s, &info);
        if (rc < 0) return rc;
        if (info.kind == ORA_THERMAL) ++*thermal;
        if (info.kind == ORA_ATHERMAL) ++*athermal;
        if (last) *last = info;
        if (rc == 1) break;
    }
    return 0;
}

static int same_info(const ora_step_info *a, const ora_step_info *b)
{
    return a->kind == b->kind && a->site == b->site && a->mode == b->mode &&
           a->n_transformed == b->n_transformed && a->totime == b->totime &&
           memcmp(a->stress, b->stress, sizeof a->stress) == 0 && memcmp(a->epsavg, b->epsavg, sizeof a->epsavg) == 0;
}

int main(int argc, char **argv)
{
    const char *image = argc > 1 ? argv[1] : "oracle_check.img";
    ora_set_threads(2);
    int th = 0, ath = 0;

    /* RNG and FFT alone */
    ora_mt r;
    ora_mt_seed(&r, 5489);
    double g = 0.0;
    for (int i = 0; i < 1000; ++i) g += ora_gaussian(&r, 1.0) + ora_mt_uniform_pos(&r);
    if (!isfinite(g)) return fail("rng");
    double fft[2 * 4 * 8 * 6];
    for (int i = 0; i < 2 * 4 * 8 * 6; ++i) fft[i] = ora_mt_uniform(&r);
    if (ora_fft3d(fft, 4, 8, 6, -1) != 0 || ora_fft3d(fft, 4, 8, 6, 1) != 0) return fail("fft");

    /* 2-D tension with the print block; barriers lowered at a few voxels for avalanches */
    ora_params p = reference_2d();
    ora_state *s = ora_create(&p);
    if (!s) return fail("create 2-D");
    ora_set_print(s, 5, 1.0);
    double *q0 = ora_field(s, ORA_F_Q0);
    for (int v = 0; v < 1024; v += 97) q0[v] = -1.0;
    if (run(s, 300, &th, &ath, NULL)) return fail("2-D steps");
    ora_softening(s);
    (void)ora_act_energy(s);
    if (!(ora_total_rate(s) > 0)) return fail("total rate");
    int site = -1, mode = -1;
    if (ora_select(s, 0.5, &site, &mode) != 0 || site < 0 || mode < 0) return fail("select");
    ora_transform(s, mode, site);
    ora_stress_solve(s);
    if (!isfinite(ora_elastic_energy(s))) return fail("energy");
    double *u = malloc(sizeof(double) * 2 * 1024);
    ora_displacement(s, u);
    free(u);
    ora_destroy(s);

    /* 3-D shear with the loading-consistent correction, heat conduction, and a state
       image written after 100 steps: a fresh state loaded from it continues identically */
    p = reference_2d();
    p.dim = 3;
    p.n[0] = 16; p.n[1] = 8; p.n[2] = 16;
    p.loading = 1;
    p.athermal_correction = 1;
    p.E = 94.5; p.nu = 0.35; p.double_nu = 0;
    p.gamma0 = 0.1; p.q0_std = 0.3; p.seed = 1234;
    s = ora_create(&p);
    if (!s) return fail("create 3-D");
    ora_set_print(s, 5, 1.0);
    ora_set_heat(s, 1);
    if (run(s, 100, &th, &ath, NULL)) return fail("3-D steps");
    if (ora_save(s, image) != 0) return fail("save");
    ora_state *t = ora_create_empty(&p);
    if (!t || ora_load(t, image) != 0) return fail("load");
    ora_step_info a, b;
    int th2 = 0, ath2 = 0;
    for (int i = 0; i < 60; ++i) {
        if (run(s, 1, &th, &ath, &a) || run(t, 1, &th2, &ath2, &b)) return fail("3-D steps after the image");
        if (!same_info(&a, &b)) return fail("image continuation differs");
    }
    if (ora_heat_conduct(s, 1e-3) != 0) return fail("heat");
    ora_destroy(t);
    ora_destroy(s);
    remove(image);

    /* per-voxel moduli: the fixed-point loop */
    p = reference_2d();
    s = ora_create(&p);
    if (!s) return fail("create moduli");
    double *lam = malloc(sizeof(double) * 1024), *mu = malloc(sizeof(double) * 1024);
    for (int v = 0; v < 1024; ++v) {
        mu[v] = 20.0 + 10.0 * ora_mt_uniform(&r);
        lam[v] = 40.0 + 10.0 * ora_mt_uniform(&r);
    }
    ora_set_moduli(s, lam, mu);
    ora_stress_solve(s);
    if (ora_inhom_iterations(s) < 1) return fail("moduli iterations");
    if (run(s, 30, &th, &ath, NULL)) return fail("moduli steps");
    ora_set_moduli(s, NULL, NULL);
    free(lam);
    free(mu);
    ora_destroy(s);

    /* crystal-like MRO, random map and a given map */
    p = reference_2d();
    p.mro_fraction = 0.3;
    s = ora_create(&p);
    if (!s) return fail("create MRO");
    ora_set_print(s, 5, 1.0);
    if (run(s, 200, &th, &ath, NULL)) return fail("MRO steps");
    const unsigned char *map = ora_mro_map(s);
    int *type = malloc(sizeof(int) * 1024);
    double *theta = malloc(sizeof(double) * 1024);
    for (int v = 0; v < 1024; ++v) {
        type[v] = map[v] ? 1 + v % 3 : 0;
        theta[v] = ora_mt_uniform(&r) - 0.5;
    }
    ora_destroy(s);
    p.mro_fraction = 0.0;
    s = ora_create_mro(&p, type, theta);
    if (!s) return fail("create MRO map");
    if (run(s, 100, &th, &ath, NULL)) return fail("MRO map steps");
    free(type);
    free(theta);
    ora_destroy(s);

    if (th < 50 || ath < 1) return fail("too few thermal or athermal steps");
    printf("ok: %d thermal and %d athermal steps\n", th, ath);
    return 0;
}
