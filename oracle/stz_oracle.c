#define _GNU_SOURCE
/*
 * stz_oracle.c -- TEST INFRASTRUCTURE ONLY (see stz_oracle.h).
 *
 * A plain-C, FP64, single-threaded restatement of the reference STZ kinetic
 * Monte Carlo step.  It follows the reference's algorithm and expression
 * order; where the reference is 2-D only, the 3-D form is the same formula
 * with six tensor components (xx, yy, zz, yz, xz, xy), contraction weights
 * (1,1,1,2,2,2) and a traceless transformation strain with five stored
 * components (xx, yy, yz, xz, xy; zz = -(xx + yy)).
 *
 * Function -> reference (STZ/ = "Mesoscale simulation STZ model/") -> pin:
 *   ora_mt_*          GSL mt19937 (seeded as STZ/init.c:142-150)
 *                     pin: std::mt19937 known answer (10000th draw of seed
 *                     5489 = 4123659995) and GSL's seed-0 -> 4357 rule,
 *                     checked in tests/test_oracle_rng.py.
 *   ora_gaussian      polar method, STZ/init.c:207-223 (= gsl_ran_gaussian)
 *                     pin: re-derived from the uniform stream in the test, plus
 *                     moments of 2e5 draws.
 *   ora_fft3d         FFTW 2.x fftwnd_mpi, unnormalised, STZ/init.c:80-83,
 *                     STZ/elastic.c:794,819 -- own Stockham mixed-radix FFT.
 *                     pin: numpy.fft golden vectors and an O(N^2) DFT.
 *   grad_tables       Grad, STZ/init.c:112-140 (incl. round() Nyquist sign,
 *                     truncated PI, g2 = 1 at k = 0).
 *                     pin: closed form k = 2*pi*(j/N - round(j/N)) in tests.
 *   hom_solve         HomogeneousStressSolver uflag=0, STZ/elastic.c:779-828,
 *                     helpers :652-747.
 *   ora_stress_solve  InhomogeneousStressSolver, STZ/elastic.c:977-1066 with
 *                     uniform moduli (STZ/init.c:457-465): one inner solve on
 *                     the round-tripped field C0^-1(C0:eps0).
 *                     pins: uniform eps0 -> sigma = C0:(epsavg - eps0);
 *                     compatible eps0 -> zero fluctuation stress; laminate
 *                     closed form; k.sigma_hat = 0 (equilibrium); brute-force
 *                     dense Green's-operator solve on a 4x4 / 4x4x4 grid.
 *   ora_heat_conduct  HeatConducting, STZ/elastic.c:549-582 (HEAT_CONDUCT; Tlocal is heated by
 *                     every transform, STZ/elastic.c:381-383, flag or not).  pins: a single
 *                     Fourier mode decays by exactly exp(-|k| dt); the k = 0 mode by
 *                     exp(-dt) (g2 = 1 there); the transform's heating against a
 *                     hand-computed sigma:deps.
 *   ora_set_moduli    Lambda(x), Mu(x) of STZ/init.c:457-465 (uniform when unset); the
 *                     wrapper then iterates to its fixed point, STZ/elastic.c:990-1062.
 *                     pins: explicit uniform moduli give the uniform path bit for bit;
 *                     a laminate under applied shear converges to the closed form
 *                     (uniform traction, harmonic-mean shear modulus).
 *   ora_displacement  print-step InhomogeneousStressSolver(uflag=2), STZ/bmg.c:289-295,
 *                     STZ/elastic.c:873-900 + helpers :754-777.
 *                     pins: the symmetric spectral gradient of u equals the solved
 *                     strain fluctuation on every non-Nyquist k; eps0 = 0 leaves the
 *                     position term alone (closed form).
 *   ora_softening     SofteningFactor default branch, STZ/init.c:711-739.
 *   ora_act_energy    ActEnergy with SF_SQUARE_FWD, MICRO_MEMORY,
 *                     NO_SF_XTAL_MRO, STZ/init.c:561-690; Q<0 count
 *                     STZ/bmg.c:147-153.  pins: closed-form single voxel; pure shear
 *                     parallel to the mode strain (W = s|deps|^2 V 1e-18/e) with last
 *                     strains parallel (D = 1), antiparallel (D = 0) and orthogonal
 *                     (D = 1/4) (tests/test_oracle_step_pins.py).
 *   ora_total_rate    KMC rate loop, STZ/evolution.c:120-134.
 *   ora_select        KMC normalise / prefix / search, STZ/evolution.c:136-224
 *                     pin: numpy cumsum + searchsorted brute force.
 *   ora_transform     VoxelShearTransform glass branch, STZ/elastic.c:361-520.
 *                     pin: hand-computed 2x2 grid, M = 2 case.
 *   ora_elastic_energy ElasticEnergy, STZ/elastic.c:533-547.  pin: sum of
 *                     sigma:C0^-1:sigma/2 with a numerical inverse of the Voigt stiffness.
 *   athermal          athermal_plasticity, STZ/evolution.c:47-100.
 *                     pin: hand-computed 2x2 grid, M = 2: first-index argmin ties, the
 *                     0.33 boundary, row-major redraw offsets in the stream, the solve's
 *                     mean C0:(epsavg - <eps0>) and the correction against a numerical
 *                     inverse of the Voigt stiffness.
 *   ora_step          main loop body, STZ/bmg.c:138-411, KMC loading
 *                     STZ/evolution.c:262-358, print block STZ/bmg.c:181-380 (Q0 <- Q1
 *                     and the re-solve), termination STZ/bmg.c:382-410.
 *                     pins: thermal step from a uniform state (R, t = 1/R, the dice's
 *                     index, mixed strain/stress loading: a fresh solve of the new
 *                     state keeps the loaded <sigma> and has zero mean stress across
 *                     the load); elastic step (dt, uniaxial-stress increment, the shear
 *                     branch's strainrate*1); the print schedule for periods 5, 3, 1,
 *                     0, -2; end-strain, tension and shear instability exits; a
 *                     crystal-like redraw undone by the next print.  Each of eight
 *                     planted mistakes in these branches fails one of the pins.
 *
 * Threads: element-wise loops and FFT lines run under OpenMP; every floating-point
 * reduction (rate total, prefix, <sigma>) stays sequential in the reference order, so
 * results are bit-identical for any thread count (tests/test_oracle_kmc.py checks it).
 *
 * Deliberate restatement choices (each observable-equivalent):
 *   - PreDEps0[M][C][N] + PreSelectMode[N] are stored as deps_last[S][N]:
 *     the reference only ever reads PreDEps0[PreSelectMode[v]][.][v]
 *     (STZ/init.c:573,613-618) and sets PreSelectMode[v] to the transformed
 *     mode right after every snapshot (STZ/evolution.c:74,207,220).
 *   - the derived components (yy = -xx in 2-D) are rebuilt on use; negation
 *     is exact, so values are bit-identical.
 *   - the O(N) dead average loop in VoxelShearTransform (STZ/elastic.c:438-448),
 *     Work[], PreQ0, Tlocal and the per-step ElasticEnergy off print steps
 *     are skipped: none of them is read.
 *   - the binary search (STZ/evolution.c:170-190) is stated as "first i with
 *     S_i >= dice"; the reference's loop has that result whenever it
 *     terminates.
 */
#include "stz_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* loops over fewer elements than this run on one thread: on the tiny grids of the pin
   tests a parallel region costs more than it saves (results are the same either way) */
#define ORA_PAR_MIN 32768L

/* ------------------------------------------------------------------ */
/* constants of STZ/bmg.h:42,68-74                                       */
/* ------------------------------------------------------------------ */
#define EV_IN_J 1.60217733e-19
#define J_IN_EV (1 / EV_IN_J)
#define ORA_EPSILON 1e-17
#define BOLZ 1.380658e-23
#define ORA_PI 3.14159265359
#define ORA_TWO_PI (2 * ORA_PI)
/* thermal_coeff1 = 1/Cp/rho with Cp = 420 J/kg-K, rho = 6125 kg/m^3; thermal_diffusivity
   3e-6 m^2/s (STZ/bmg.h:185-188) */
#define ORA_THERMAL_COEFF1 (1 / 420.0 / 6125.0)
#define ORA_THERMAL_DIFFUSIVITY 3E-6

/* ------------------------------------------------------------------ */
/* MT19937 as GSL's mt19937 (seed 0 -> 4357)                             */
/* ------------------------------------------------------------------ */
void ora_mt_seed(ora_mt *r, unsigned long s)
{
    if (s == 0) s = 4357;
    r->mt[0] = (uint32_t)(s & 0xffffffffUL);
    for (int i = 1; i < 624; i++)
        r->mt[i] = (uint32_t)(1812433253UL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (unsigned long)i);
    r->mti = 624;
}

uint32_t ora_mt_next(ora_mt *r)
{
    static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
    uint32_t k;
    if (r->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; kk++) {
            uint32_t y = (r->mt[kk] & 0x80000000u) | (r->mt[kk + 1] & 0x7fffffffu);
            r->mt[kk] = r->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < 623; kk++) {
            uint32_t y = (r->mt[kk] & 0x80000000u) | (r->mt[kk + 1] & 0x7fffffffu);
            r->mt[kk] = r->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        {
            uint32_t y = (r->mt[623] & 0x80000000u) | (r->mt[0] & 0x7fffffffu);
            r->mt[623] = r->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        }
        r->mti = 0;
    }
    k = r->mt[r->mti++];
    k ^= (k >> 11);
    k ^= (k << 7) & 0x9d2c5680u;
    k ^= (k << 15) & 0xefc60000u;
    k ^= (k >> 18);
    return k;
}

double ora_mt_uniform(ora_mt *r) { return ora_mt_next(r) / 4294967296.0; }

double ora_mt_uniform_pos(ora_mt *r)
{
    double x;
    do { x = ora_mt_uniform(r); } while (x == 0);
    return x;
}

/* STZ/init.c:207-223 */
double ora_gaussian(ora_mt *r, double sigma)
{
    double x, y, r2;
    do {
        x = -1 + 2 * ora_mt_uniform_pos(r);
        y = -1 + 2 * ora_mt_uniform_pos(r);
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0);
    return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

/* ------------------------------------------------------------------ */
/* FFT: Stockham autosort, mixed radix, twiddles from long double        */
/* ------------------------------------------------------------------ */
typedef struct { int n, nf, fac[40]; double *w; } plan1d;

static int plan_init(plan1d *p, int n)
{
    int m = n;
    p->n = n; p->nf = 0;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
    p->w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!p->w) return -1;
    for (int t = 0; t < n; t++) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        p->w[2 * t] = (double)cosl(a);
        p->w[2 * t + 1] = -(double)sinl(a); /* e^{-2 pi i t / n} */
    }
    return 0;
}

static void plan_free(plan1d *p) { free(p->w); p->w = NULL; }

/* x (n complex) -> x, y scratch; sign -1 forward, +1 backward */
static void fft1d(const plan1d *p, double *x, double *y, int sign)
{
    const int n = p->n;
    double *src = x, *dst = y;
    int Ns = 1;
    double v[2 * 64], o[2 * 64];
    for (int f = 0; f < p->nf; f++) {
        const int q = p->fac[f];
        const int nq = n / q;
        const int tstep = n / (Ns * q);
        for (int j = 0; j < nq; j++) {
            const int jm = j % Ns;
            for (int r = 0; r < q; r++) {
                double re = src[2 * (j + r * nq)], im = src[2 * (j + r * nq) + 1];
                int t = (jm * tstep * r) % n;
                double wr = p->w[2 * t], wi = sign < 0 ? p->w[2 * t + 1] : -p->w[2 * t + 1];
                v[2 * r] = re * wr - im * wi;
                v[2 * r + 1] = re * wi + im * wr;
            }
            if (q == 2) {
                o[0] = v[0] + v[2]; o[1] = v[1] + v[3];
                o[2] = v[0] - v[2]; o[3] = v[1] - v[3];
            } else if (q == 4) {
                double s = (double)sign;
                double acr = v[0] + v[4], aci = v[1] + v[5];
                double amr = v[0] - v[4], ami = v[1] - v[5];
                double bdr = v[2] + v[6], bdi = v[3] + v[7];
                double bmr = v[2] - v[6], bmi = v[3] - v[7];
                o[0] = acr + bdr; o[1] = aci + bdi;
                o[4] = acr - bdr; o[5] = aci - bdi;
                /* s*i*(bm) = (-s*bmi, s*bmr) */
                o[2] = amr - s * bmi; o[3] = ami + s * bmr;
                o[6] = amr + s * bmi; o[7] = ami - s * bmr;
            } else {
                for (int k = 0; k < q; k++) {
                    double ar = 0, ai = 0;
                    for (int r = 0; r < q; r++) {
                        int t = ((r * k) % q) * nq;
                        double wr = p->w[2 * t], wi = sign < 0 ? p->w[2 * t + 1] : -p->w[2 * t + 1];
                        ar += v[2 * r] * wr - v[2 * r + 1] * wi;
                        ai += v[2 * r] * wi + v[2 * r + 1] * wr;
                    }
                    o[2 * k] = ar; o[2 * k + 1] = ai;
                }
            }
            const int idxD = (j / Ns) * Ns * q + jm;
            for (int r = 0; r < q; r++) {
                dst[2 * (idxD + r * Ns)] = o[2 * r];
                dst[2 * (idxD + r * Ns) + 1] = o[2 * r + 1];
            }
        }
        double *t = src; src = dst; dst = t;
        Ns *= q;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

int ora_fft3d(double *data, int n0, int n1, int n2, int sign)
{
    const int nn[3] = {n0, n1, n2};
    const long strides[3] = {1, n0, (long)n0 * n1};
    const long ntot = (long)n0 * n1 * n2;
    for (int ax = 0; ax < 3; ax++) {
        const int n = nn[ax];
        if (n == 1) continue;
        plan1d p;
        if (plan_init(&p, n)) return -1;
        const long st = strides[ax];
        const long nlines = ntot / n;
        int bad = 0;
#pragma omp parallel if(ntot >= ORA_PAR_MIN)
        {
            double *line = (double *)malloc(sizeof(double) * 4 * (size_t)n);
            if (!line) {
#pragma omp atomic write
                bad = 1;
            } else {
                double *work = line + 2 * n;
#pragma omp for schedule(static)
                for (long li = 0; li < nlines; li++) {
                    /* line li: index 0 along ax; lower axes fastest, then higher */
                    const long lo = li % st, hi = li / st;
                    const long base = hi * st * n + lo;
                    for (int i = 0; i < n; i++) {
                        line[2 * i] = data[2 * (base + i * st)];
                        line[2 * i + 1] = data[2 * (base + i * st) + 1];
                    }
                    fft1d(&p, line, work, sign);
                    for (int i = 0; i < n; i++) {
                        data[2 * (base + i * st)] = line[2 * i];
                        data[2 * (base + i * st) + 1] = line[2 * i + 1];
                    }
                }
                free(line);
            }
        }
        plan_free(&p);
        if (bad) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* state                                                                */
/* ------------------------------------------------------------------ */
struct ora_state {
    ora_params p;
    int C, S, M;
    long N;
    int lc;                 /* loading component: 0 (tension) or xy */
    /* derived constants (STZ/init.c:6-75) */
    double E, nu, mu, lambda, kT, tTemp, dt, voxel, ratepre;
    double permsoft, tempsoft, permcap;
    /* k tables (STZ/init.c:112-140) */
    double *g[3], *gg[6];
    /* fields */
    double *eps0, *sigma, *epso, *q0, *deps, *deps_last, *perm, *temp, *tlast, *q, *soft;
    double *q1;             /* stress-free barriers, restored at every print (STZ/init.c:189, STZ/bmg.c:215-223) */
    double *rate;
    /* solver scratch */
    double *spec;           /* C complex fields, N each */
    double *tau0, *epsolast;
    /* crystal-like MRO: type per voxel (0 glass, 1 six-fold, 2 two-fold, 3 ten-fold) and
       the orientation MROthe0 (STZ/init.c:345-437) */
    unsigned char *mro;
    double *mroth;
    /* local temperature (STZ/init.c:470-476) and |k| of the reference grid (Grad's g2) */
    double *tlocal, *g2;
    int heat;               /* HEAT_CONDUCT on */
    /* elastic inhomogeneity: Lambda[N], Mu[N] (NULL: uniform, STZ/init.c:457-465) */
    double *lamx, *mux, *estar;
    int inhom_iters;
    /* scalars */
    double epsavg[6], stress[6], totime;
    ora_mt rng;
    long long step, astep, kstep, estep;
    long long avlan_counter;
    double soc_elast;
    int print_on, printperiod; double endstrain; double strainhist_last;
};

static const double W3[3] = {1, 1, 2};
static const double W6[6] = {1, 1, 1, 2, 2, 2};

static inline const double *wts(const ora_state *s) { return s->C == 3 ? W3 : W6; }

/* a:b with weights, left-to-right as STZ/init.c:585-586 */
static inline double contract(const ora_state *s, const double *a, const double *b)
{
    const double *w = wts(s);
    double acc = a[0] * b[0];
    for (int c = 1; c < s->C; c++)
        acc = acc + (w[c] == 2 ? (2 * a[c]) * b[c] : a[c] * b[c]);
    return acc;
}

/* stored S components -> C tensor components */
static inline void expand(const ora_state *s, const double *st, double *t)
{
    if (s->C == 3) { t[0] = st[0]; t[1] = -st[0]; t[2] = st[1]; }
    else { t[0] = st[0]; t[1] = st[1]; t[2] = -(st[0] + st[1]); t[3] = st[2]; t[4] = st[3]; t[5] = st[4]; }
}

static inline void deps_of(const ora_state *s, int m, long v, double *t)
{
    double st[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < s->S; k++) st[k] = s->deps[((long)m * s->S + k) * s->N + v];
    expand(s, st, t);
}

static inline void deps_last_of(const ora_state *s, long v, double *t)
{
    double st[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < s->S; k++) st[k] = s->deps_last[(long)k * s->N + v];
    expand(s, st, t);
}

/* C0 : e (IsotropicMul_RF / IsotropicMul_C per real part), STZ/elastic.c:905-920 */
static inline void iso_mul(const ora_state *s, double lam, double mu, const double *e, double *t)
{
    if (s->C == 3) {
        double tra = lam * (e[0] + e[1]);
        t[0] = tra + 2 * mu * e[0];
        t[1] = tra + 2 * mu * e[1];
        t[2] = 2 * mu * e[2];
    } else {
        double tra = lam * ((e[0] + e[1]) + e[2]);
        for (int c = 0; c < 3; c++) t[c] = tra + 2 * mu * e[c];
        for (int c = 3; c < 6; c++) t[c] = 2 * mu * e[c];
    }
}

/* C0^-1 : t (IsotropicInv), STZ/elastic.c:940-955 */
static inline void iso_inv(const ora_state *s, double lam, double mu, const double *t, double *e)
{
    if (s->C == 3) {
        double tra = -1.0 * lam / 2.0 / mu / (2.0 * lam + 2.0 * mu) * (t[0] + t[1]);
        e[0] = tra + t[0] / 2.0 / mu;
        e[1] = tra + t[1] / 2.0 / mu;
        e[2] = t[2] / 2.0 / mu;
    } else {
        double tra = -1.0 * lam / 2.0 / mu / (3.0 * lam + 2.0 * mu) * ((t[0] + t[1]) + t[2]);
        for (int c = 0; c < 3; c++) e[c] = tra + t[c] / 2.0 / mu;
        for (int c = 3; c < 6; c++) e[c] = t[c] / 2.0 / mu;
    }
}

/* isotropicinv, STZ/elastic.c:522-531 (scalar form used by the loading) */
static inline void isotropicinv(const ora_state *s, const double *st, double *e)
{
    const double lam = s->lambda, mu = s->mu;
    if (s->C == 3) {
        double tra = -lam / 2 / mu / (2 * lam + 2 * mu) * (st[0] + st[1]);
        e[0] = tra + st[0] / 2 / mu;
        e[1] = tra + st[1] / 2 / mu;
        e[2] = st[2] / 2 / mu;
    } else {
        double tra = -lam / 2 / mu / (3 * lam + 2 * mu) * ((st[0] + st[1]) + st[2]);
        for (int c = 0; c < 3; c++) e[c] = tra + st[c] / 2 / mu;
        for (int c = 3; c < 6; c++) e[c] = st[c] / 2 / mu;
    }
}

static void grad_tables(ora_state *s)
{
    const int *n = s->p.n;
    const long N = s->N;
    for (int a = 0; a < 3; a++) s->g[a] = (double *)calloc((size_t)N, sizeof(double));
    for (int a = 0; a < 6; a++) s->gg[a] = (double *)calloc((size_t)N, sizeof(double));
    for (long v = 0; v < N; v++) {
        int idx[3] = {(int)(v % n[0]), (int)((v / n[0]) % n[1]), (int)(v / ((long)n[0] * n[1]))};
        double g[3] = {0, 0, 0};
        for (int a = 0; a < s->p.dim; a++) {
            double sa = (double)(idx[a]) * (double)(1.0 / n[a]);
            g[a] = (sa - round(sa)) * ORA_TWO_PI;
        }
        double g2;
        if (s->p.dim == 2) g2 = sqrt(g[0] * g[0] + g[1] * g[1]);
        else g2 = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (idx[0] + idx[1] + idx[2] == 0) g2 = 1.0;
        s->g2[v] = g2;
        for (int a = 0; a < 3; a++) { g[a] /= g2; s->g[a][v] = g[a]; }
        if (s->p.dim == 2) {
            s->gg[0][v] = g[0] * g[0];
            s->gg[1][v] = g[1] * g[1];
            s->gg[2][v] = g[0] * g[1];
        } else {
            s->gg[0][v] = g[0] * g[0];
            s->gg[1][v] = g[1] * g[1];
            s->gg[2][v] = g[2] * g[2];
            s->gg[3][v] = g[1] * g[2];
            s->gg[4][v] = g[0] * g[2];
            s->gg[5][v] = g[0] * g[1];
        }
    }
}

static void derive(ora_state *s)
{
    const ora_params *p = &s->p;
    s->E = p->E;
    s->nu = p->nu;
    if (p->double_nu) s->nu *= 2;                         /* STZ/init.c:26 */
    s->mu = s->E / (2 + 2 * s->nu);                        /* STZ/init.c:27 */
    if (p->dim == 2) s->lambda = s->E * s->nu / (1 - s->nu * s->nu); /* STZ/init.c:44 */
    else s->lambda = s->E * s->nu / ((1 + s->nu) * (1 - 2 * s->nu));
    s->kT = p->temperature * BOLZ * J_IN_EV;               /* STZ/bmg.h:73 */
    s->dt = p->strainincrement / p->strainrate;            /* STZ/init.c:48 */
    s->tTemp = 1 / (p->trial_frequency * exp(-p->temp_soft_barrier / s->kT)); /* STZ/init.c:49 */
    s->voxel = p->lx * p->ly * p->lz;                      /* STZ/bmg.h:90 */
    s->ratepre = s->voxel * p->trial_frequency;            /* STZ/evolution.c:125 */
    s->permsoft = p->perm_soft;
    s->tempsoft = p->temp_soft;
    s->permcap = p->perm_cap;
}

static ora_state *alloc_state(const ora_params *p)
{
    if (!p || (p->dim != 2 && p->dim != 3) || p->modes <= 0) return NULL;
    ora_state *s = (ora_state *)calloc(1, sizeof(ora_state));
    if (!s) return NULL;
    s->p = *p;
    if (p->dim == 2) s->p.n[2] = 1;
    s->C = p->dim == 2 ? 3 : 6;
    s->S = p->dim == 2 ? 2 : 5;
    s->M = p->modes;
    s->N = (long)s->p.n[0] * s->p.n[1] * s->p.n[2];
    s->lc = p->loading == 0 ? 0 : (s->C == 3 ? 2 : 5);
    derive(s);
    const long N = s->N, C = s->C, S = s->S, M = s->M;
    s->eps0 = (double *)calloc((size_t)(C * N), sizeof(double));
    s->sigma = (double *)calloc((size_t)(C * N), sizeof(double));
    s->epso = (double *)calloc((size_t)(C * N), sizeof(double));
    s->q0 = (double *)calloc((size_t)(M * N), sizeof(double));
    s->q1 = (double *)calloc((size_t)(M * N), sizeof(double));
    s->deps = (double *)calloc((size_t)(M * S * N), sizeof(double));
    s->deps_last = (double *)calloc((size_t)(S * N), sizeof(double));
    s->perm = (double *)calloc((size_t)N, sizeof(double));
    s->temp = (double *)calloc((size_t)N, sizeof(double));
    s->tlast = (double *)calloc((size_t)N, sizeof(double));
    s->q = (double *)calloc((size_t)(M * N), sizeof(double));
    s->soft = (double *)calloc((size_t)N, sizeof(double));
    s->rate = (double *)calloc((size_t)(M * N), sizeof(double));
    s->spec = (double *)calloc((size_t)(2 * C * N), sizeof(double));
    s->tau0 = (double *)calloc((size_t)(C * N), sizeof(double));
    s->epsolast = (double *)calloc((size_t)(C * N), sizeof(double));
    s->tlocal = (double *)calloc((size_t)N, sizeof(double));
    s->g2 = (double *)calloc((size_t)N, sizeof(double));
    s->mro = (unsigned char *)calloc((size_t)N, 1);
    s->mroth = (double *)calloc((size_t)N, sizeof(double));
    if (!s->mro || !s->mroth || !s->eps0 || !s->sigma || !s->epso || !s->q0 || !s->q1 || !s->deps || !s->deps_last || !s->perm ||
        !s->temp || !s->tlast || !s->q || !s->soft || !s->rate || !s->spec || !s->tau0 || !s->epsolast ||
        !s->tlocal || !s->g2) {
        ora_destroy(s);
        return NULL;
    }
    grad_tables(s);
    for (long v = 0; v < N; v++) s->tlast[v] = -100 * s->tTemp; /* STZ/init.c:312 */
    for (long v = 0; v < N; v++) s->tlocal[v] = p->temperature; /* STZ/init.c:470-476 */
    ora_mt_seed(&s->rng, (unsigned long)p->seed);              /* STZ/init.c:143-147, rank 0 */
    s->printperiod = 0;
    s->endstrain = 1e300;
    return s;
}

ora_state *ora_create_empty(const ora_params *p) { return alloc_state(p); }

void ora_destroy(ora_state *s)
{
    if (!s) return;
    free(s->eps0); free(s->sigma); free(s->epso); free(s->q0); free(s->q1); free(s->deps); free(s->deps_last);
    free(s->perm); free(s->temp); free(s->tlast); free(s->q); free(s->soft); free(s->rate);
    free(s->spec); free(s->tau0); free(s->epsolast);
    free(s->lamx); free(s->mux); free(s->estar);
    free(s->tlocal); free(s->g2);
    free(s->mro); free(s->mroth);
    for (int a = 0; a < 3; a++) free(s->g[a]);
    for (int a = 0; a < 6; a++) free(s->gg[a]);
    free(s);
}

/* ------------------------------------------------------------------ */
/* state image (test infrastructure: lets a long golden run resume)     */
/* ------------------------------------------------------------------ */
#define ORA_IMAGE_MAGIC 0x4f524131u   /* "ORA1" */

/* every array of the state with its length in bytes (NULL arrays have length 0) */
static int state_arrays(ora_state *s, void **ptr[32], size_t len[32])
{
    const size_t N = (size_t)s->N, C = (size_t)s->C, S = (size_t)s->S, M = (size_t)s->M, d = sizeof(double);
    int k = 0;
#define ORA_ARR(member, bytes) do { ptr[k] = (void **)&s->member; len[k] = s->member ? (bytes) : 0; k++; } while (0)
    ORA_ARR(eps0, C * N * d); ORA_ARR(sigma, C * N * d); ORA_ARR(epso, C * N * d);
    ORA_ARR(q0, M * N * d); ORA_ARR(q1, M * N * d); ORA_ARR(deps, M * S * N * d); ORA_ARR(deps_last, S * N * d);
    ORA_ARR(perm, N * d); ORA_ARR(temp, N * d); ORA_ARR(tlast, N * d); ORA_ARR(q, M * N * d);
    ORA_ARR(soft, N * d); ORA_ARR(rate, M * N * d); ORA_ARR(spec, 2 * C * N * d); ORA_ARR(tau0, C * N * d);
    ORA_ARR(epsolast, C * N * d); ORA_ARR(tlocal, N * d); ORA_ARR(g2, N * d); ORA_ARR(mro, N);
    ORA_ARR(mroth, N * d); ORA_ARR(lamx, N * d); ORA_ARR(mux, N * d); ORA_ARR(estar, C * N * d);
    for (int a = 0; a < 3; a++) ORA_ARR(g[a], N * d);
    for (int a = 0; a < 6; a++) ORA_ARR(gg[a], N * d);
#undef ORA_ARR
    return k;
}

int ora_save(ora_state *s, const char *path)
{
    if (!s || !path) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    void **ptr[32];
    size_t len[32];
    const int na = state_arrays(s, ptr, len);
    const unsigned magic = ORA_IMAGE_MAGIC;
    int ok = fwrite(&magic, sizeof magic, 1, f) == 1 && fwrite(s, sizeof *s, 1, f) == 1 &&
             fwrite(len, sizeof(size_t), (size_t)na, f) == (size_t)na;
    for (int i = 0; ok && i < na; i++)
        if (len[i]) ok = fwrite(*ptr[i], 1, len[i], f) == len[i];
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -3;
}

/* s must have been made with the same parameters (and set_moduli if the image has
   per-voxel moduli); every scalar and array of the image replaces s's own */
int ora_load(ora_state *s, const char *path)
{
    if (!s || !path) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    unsigned magic = 0;
    ora_state img;
    void **ptr[32];
    size_t len[32], want[32];
    const int na = state_arrays(s, ptr, want);
    int ok = fread(&magic, sizeof magic, 1, f) == 1 && magic == ORA_IMAGE_MAGIC && fread(&img, sizeof img, 1, f) == 1 &&
             fread(len, sizeof(size_t), (size_t)na, f) == (size_t)na;
    ok = ok && img.N == s->N && img.C == s->C && img.M == s->M && img.S == s->S;
    for (int i = 0; ok && i < na; i++) ok = len[i] == want[i];
    if (ok) {
        /* the image's scalars with s's own arrays */
        void *keep[32];
        for (int i = 0; i < na; i++) keep[i] = *ptr[i];
        *s = img;
        for (int i = 0; i < na; i++) *ptr[i] = keep[i];
        for (int i = 0; ok && i < na; i++)
            if (len[i]) ok = fread(*ptr[i], 1, len[i], f) == len[i];
    }
    fclose(f);
    return ok ? 0 : -3;
}

static void draw_deps(ora_state *s, int m, long v)
{
    const double a = 0.5 * s->p.gamma0;   /* 0.5*GAMMA0 folded first, STZ/elastic.c:127 */
    for (int k = 0; k < s->S; k++)
        s->deps[((long)m * s->S + k) * s->N + v] = a * ora_gaussian(&s->rng, 1.0);
}

/* modes of a crystal-like MRO family and the scale of its mode-0 strain
   (STZ/bmg.h:30-32; STZ/elastic.c:67-69, 86-88, 105-107) */
static inline int mro_modes(int t) { return t == 1 ? 6 : (t == 2 ? 2 : 12); }

/* DStrain for one (mode, voxel), STZ/elastic.c:60-138: glass draws two Gaussians; a
   crystal-like voxel's mode 0 is the family's strain along MROthe0, modes 1..XMODES-1 its
   rotations by 2*theta_m, theta_m = m*2*PI/XMODES*(1 + 0.2(2U - 1)), each term scaled by
   (1 + 0.2(2U - 1)); modes from XMODES on have zero strain.  The four uniforms inside the
   two expressions are drawn left operand first, the order g++ -O3 -std=c++11 (the
   reference's build, STZ/Makefile:1-3) evaluates them in. */
static void dstrain(ora_state *s, int m, long v)
{
    const int t = s->mro[v];
    if (t == 0) { draw_deps(s, m, v); return; }
    const long N = s->N;
    double *xx = &s->deps[((long)m * s->S + 0) * N + v];
    double *xy = &s->deps[((long)m * s->S + 1) * N + v];
    const int XM = mro_modes(t);
    if (m == 0) {
        const double amp = t == 2 ? 0.8 * s->p.gamma0 : 0.5 * s->p.gamma0;
        /* two separate calls: the reference reads MROthe0[idx] again after storing through
           the DEps0 pointers, so the compiler cannot pair them */
        *xx = amp * cos(s->mroth[v]);
        *xy = amp * sin(s->mroth[v]);
    } else if (m < XM) {
        const double a0 = s->deps[(0 + 0) * N + v], a2 = s->deps[(0 + 1) * N + v];
        const double th = m * 2. * ORA_PI / XM * (1. + 0.2 * (2 * ora_mt_uniform(&s->rng) - 1.));
        /* cos(2*th) and sin(2*th) of one local: g++ computes the pair with one sincos call,
           whose sine can differ from sin() in the last bit */
        double s2, c2;
        sincos(2 * th, &s2, &c2);
        const double u1 = ora_mt_uniform(&s->rng);
        const double u2 = ora_mt_uniform(&s->rng);
        *xx = a0 * (1. + 0.2 * (2 * u1 - 1.)) * c2 + a2 * (1. + 0.2 * (2 * u2 - 1.)) * s2;
        const double u3 = ora_mt_uniform(&s->rng);
        const double u4 = ora_mt_uniform(&s->rng);
        *xy = -a0 * (1. + 0.2 * (2 * u3 - 1.)) * s2 + a2 * (1. + 0.2 * (2 * u4 - 1.)) * c2;
    } else {
        *xx = 0.;
        *xy = 0.;
    }
}

/* Energy_Init (STZ/init.c:181-198, every voxel alike), DStrain, initial solve */
static ora_state *init_fields(ora_state *s)
{
    const long N = s->N;
    for (long v = 0; v < N; v++)
        if (s->mro[v] && s->p.dim == 3) { ora_destroy(s); return NULL; }   /* 2-D families only */
    for (int m = 0; m < s->M; m++)
        for (long v = 0; v < N; v++)
            s->q0[(long)m * N + v] = ora_gaussian(&s->rng, s->p.q0_std) + s->p.q0_mean;
    memcpy(s->q1, s->q0, sizeof(double) * (size_t)s->M * (size_t)N);       /* Q1 = Q0, STZ/init.c:189 */
    for (int m = 0; m < s->M; m++)
        for (long v = 0; v < N; v++) dstrain(s, m, v);
    /* initial solve with eps0 = 0, epsavg = 0 (STZ/bmg.c:117-127) */
    for (int c = 0; c < 6; c++) s->epsavg[c] = 0.0;
    ora_stress_solve(s);
    return s;
}

ora_state *ora_create(const ora_params *p)
{
    ora_state *s = alloc_state(p);
    if (!s) return NULL;
    const long N = s->N;
    /* MRO map, InitMRO 0 (STZ/init.c:350-397): one uniform per voxel, six-fold crystal-like
       MRO when it is <= xMROFrac, with MROthe0 = U*PI - PI/2 from a second uniform */
    for (long v = 0; v < N; v++) {
        const double tmp = ora_mt_uniform(&s->rng);
        if (tmp <= s->p.mro_fraction) {
            s->mro[v] = 1;
            s->mroth[v] = ora_mt_uniform(&s->rng) * ORA_PI - ORA_PI / 2.0;
        } else {
            s->mro[v] = 0;
            s->mroth[v] = 0.;
        }
    }
    return init_fields(s);
}

ora_state *ora_create_mro(const ora_params *p, const int *type, const double *theta)
{
    ora_state *s = alloc_state(p);
    if (!s) return NULL;
    const long N = s->N;
    for (long v = 0; v < N; v++) {
        if (type[v] < 0 || type[v] > 3) { ora_destroy(s); return NULL; }   /* STZ/elastic.c:131-136 */
        s->mro[v] = (unsigned char)type[v];
        s->mroth[v] = theta[v];
    }
    /* InitMRO 1 (STZ/init.c:398-431): ten-fold voxels become two-fold with probability 0.5 */
    for (long v = 0; v < N; v++)
        if (s->mro[v] == 3) {
            const double tmp = ora_mt_uniform(&s->rng);
            if (tmp < 0.5) s->mro[v] = 2;
        }
    return init_fields(s);
}

const unsigned char *ora_mro_map(ora_state *s) { return s->mro; }

double *ora_field(ora_state *s, int which)
{
    switch (which) {
    case ORA_F_EPS0: return s->eps0;
    case ORA_F_SIGMA: return s->sigma;
    case ORA_F_Q0: return s->q0;
    case ORA_F_DEPS: return s->deps;
    case ORA_F_DEPS_LAST: return s->deps_last;
    case ORA_F_PERM: return s->perm;
    case ORA_F_TEMP: return s->temp;
    case ORA_F_TLAST: return s->tlast;
    case ORA_F_Q: return s->q;
    case ORA_F_SOFTENING: return s->soft;
    case ORA_F_EPSO: return s->epso;
    case ORA_F_TLOCAL: return s->tlocal;
    case ORA_F_Q1: return s->q1;
    default: return NULL;
    }
}

void ora_scalars(ora_state *s, double *epsavg, double *stress, double *totime)
{
    for (int c = 0; c < 6; c++) { if (epsavg) epsavg[c] = s->epsavg[c]; if (stress) stress[c] = s->stress[c]; }
    if (totime) *totime = s->totime;
}

void ora_set_scalars(ora_state *s, const double *epsavg, const double *stress, double totime)
{
    for (int c = 0; c < 6; c++) { if (epsavg) s->epsavg[c] = epsavg[c]; if (stress) s->stress[c] = stress[c]; }
    s->totime = totime;
}

void ora_consts(ora_state *s, double *o)
{
    o[0] = s->E; o[1] = s->nu; o[2] = s->mu; o[3] = s->lambda; o[4] = s->kT; o[5] = s->tTemp;
    o[6] = s->dt; o[7] = s->voxel; o[8] = s->ratepre; o[9] = (s->lambda + s->mu) / (s->lambda + 2.0 * s->mu);
    o[10] = (double)s->N; o[11] = s->C; o[12] = s->S; o[13] = s->M; o[14] = s->permcap; o[15] = s->lc;
}

ora_mt *ora_rng(ora_state *s) { return &s->rng; }

void ora_counters(ora_state *s, long long *o)
{
    o[0] = s->step; o[1] = s->astep; o[2] = s->kstep; o[3] = s->estep;
    o[4] = s->avlan_counter; o[5] = 0; o[6] = 0; o[7] = 0;
    memcpy(&o[5], &s->soc_elast, sizeof(double));
}

void ora_set_print(ora_state *s, int printperiod, double endstrain)
{
    s->print_on = 1;
    s->printperiod = printperiod;
    s->endstrain = endstrain;
}

void ora_set_strainhist(ora_state *s, double last) { s->strainhist_last = last; }

int ora_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* stress solve                                                          */
/* ------------------------------------------------------------------ */

/* HomogeneousStressSolver(uflag=0), STZ/elastic.c:779-828 on eps0o = s->epsolast */
static void hom_solve(ora_state *s)
{
    const long N = s->N;
    const int C = s->C;
    const double lambda = s->lambda, mu = s->mu;
    const double alpha = (lambda + mu) / (lambda + 2.0 * mu);   /* :786 */
    double *sp = s->spec;
    for (int c = 0; c < C; c++) {
        double *f = sp + 2 * (long)c * N;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) { f[2 * v] = s->epsolast[(long)c * N + v]; f[2 * v + 1] = 0.; }
        ora_fft3d(f, s->p.n[0], s->p.n[1], s->p.n[2], -1);
    }
    const double s0 = 1.0 / mu, sh = 0.5 / mu, sa = -1.0 * alpha / mu;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        for (int ri = 0; ri < 2; ri++) {
            double e[6] = {0}, t[6];
            for (int c = 0; c < C; c++) e[c] = sp[2 * ((long)c * N + v) + ri];
            iso_mul(s, lambda, mu, e, t);                 /* IsotropicMul_C :802 */
            double o[6];
            if (C == 3) {
                const double g0 = s->g[0][v], g1 = s->g[1][v];
                double F0 = t[0] * g0 + t[2] * g1;        /* :804 */
                double F1 = t[2] * g0 + t[1] * g1;        /* :805 */
                double G = F0 * g0 + F1 * g1;             /* :806 */
                o[0] = s0 * (F0 * g0) + sa * (G * s->gg[0][v]);          /* :807 */
                o[1] = s0 * (F1 * g1) + sa * (G * s->gg[1][v]);          /* :808 */
                o[2] = sh * (F0 * g1 + F1 * g0) + sa * (G * s->gg[2][v]); /* :809 */
            } else {
                const double g0 = s->g[0][v], g1 = s->g[1][v], g2 = s->g[2][v];
                /* Voigt (xx, yy, zz, yz, xz, xy) */
                double F0 = t[0] * g0 + t[5] * g1 + t[4] * g2;
                double F1 = t[5] * g0 + t[1] * g1 + t[3] * g2;
                double F2 = t[4] * g0 + t[3] * g1 + t[2] * g2;
                double G = F0 * g0 + F1 * g1 + F2 * g2;
                o[0] = s0 * (F0 * g0) + sa * (G * s->gg[0][v]);
                o[1] = s0 * (F1 * g1) + sa * (G * s->gg[1][v]);
                o[2] = s0 * (F2 * g2) + sa * (G * s->gg[2][v]);
                o[3] = sh * (F1 * g2 + F2 * g1) + sa * (G * s->gg[3][v]);
                o[4] = sh * (F0 * g2 + F2 * g0) + sa * (G * s->gg[4][v]);
                o[5] = sh * (F0 * g1 + F1 * g0) + sa * (G * s->gg[5][v]);
            }
            for (int c = 0; c < C; c++) sp[2 * ((long)c * N + v) + ri] = o[c];
        }
    }
    for (int c = 0; c < C; c++) {
        double *f = sp + 2 * (long)c * N;
        ora_fft3d(f, s->p.n[0], s->p.n[1], s->p.n[2], +1);
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) s->epso[(long)c * N + v] = f[2 * v] / (double)N + s->epsavg[c]; /* :822 */
    }
    /* IsotropicMul_R2, STZ/elastic.c:686-701 */
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        double d[6] = {0};
        for (int c = 0; c < C; c++) d[c] = s->epso[(long)c * N + v] - s->epsolast[(long)c * N + v];
        if (C == 3) {
            double tra = lambda * (d[0] + d[1]);
            s->sigma[v] = tra + 2.0 * mu * d[0];
            s->sigma[N + v] = tra + 2.0 * mu * d[1];
            s->sigma[2 * N + v] = 2.0 * mu * d[2];
        } else {
            double tra = lambda * ((d[0] + d[1]) + d[2]);
            for (int c = 0; c < 3; c++) s->sigma[(long)c * N + v] = tra + 2.0 * mu * d[c];
            for (int c = 3; c < 6; c++) s->sigma[(long)c * N + v] = 2.0 * mu * d[c];
        }
    }
    /* AvgStress, STZ/elastic.c:733-751 */
    for (int c = 0; c < C; c++) {
        double sum = 0.0;
        for (long v = 0; v < N; v++) sum += s->sigma[(long)c * N + v];
        s->stress[c] = sum / (double)N;
    }
}

/* InhomogeneousStressSolver with Lambda = lambda, Mu = mu (STZ/elastic.c:977-1066) */
static void stress_solve_uniform(ora_state *s)
{
    const long N = s->N;
    const int C = s->C;
    const double lambda = s->lambda, mu = s->mu;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        double e[6] = {0}, t0[6], t1[6], tau[6], z[6] = {0}, r[6];
        for (int c = 0; c < C; c++) { e[c] = s->eps0[(long)c * N + v]; z[c] = e[c] + s->epsavg[c]; }
        iso_mul(s, lambda, mu, e, t0);                         /* Tau0o = C(x):eps0, :1001 */
        /* Tau1o = (C0 - C(x)) : (eps0 + epsavg), :1002; zero for uniform moduli */
        if (C == 3) {
            double tra = (lambda - lambda) * (z[0] + z[1]);
            t1[0] = tra + 2 * (mu - mu) * z[0];
            t1[1] = tra + 2 * (mu - mu) * z[1];
            t1[2] = 2 * (mu - mu) * z[2];
        } else {
            double tra = (lambda - lambda) * ((z[0] + z[1]) + z[2]);
            for (int c = 0; c < 3; c++) t1[c] = tra + 2 * (mu - mu) * z[c];
            for (int c = 3; c < 6; c++) t1[c] = 2 * (mu - mu) * z[c];
        }
        for (int c = 0; c < C; c++) tau[c] = t0[c] + t1[c];
        iso_inv(s, lambda, mu, tau, r);                         /* :1009 */
        for (int c = 0; c < C; c++) {
            s->tau0[(long)c * N + v] = t0[c];
            s->epsolast[(long)c * N + v] = r[c];                /* :1017-1022 */
        }
    }
    hom_solve(s);
    s->inhom_iters = 1;
    /* Tau1o = 0 again -> epso = C0^-1 Tau0 == epsolast, mix 0.5 leaves it, convg = 0 -> break
       (STZ/elastic.c:1027-1058).  The stress field and <sigma> of hom_solve are final. */
}

/* The same wrapper with Lambda(x), Mu(x) set: the fixed point of the reference-medium
   eigenstrain eps* = C0^-1 (C(x):eps0 + (C0 - C(x)):eps), eps the total strain of the
   homogeneous solve of eps*, mixed 0.5/0.5, until the RMS change is below
   1e-3 |epsavg| (or |epsavg| < 1e-4), at most 101 solves (STZ/elastic.c:990-1062).
   |epsavg| and the RMS change run over all C components (the reference's three in 2-D). */
static void stress_solve_inhom(ora_state *s)
{
    const long N = s->N;
    const int C = s->C;
    const double lambda = s->lambda, mu = s->mu;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        double e[6] = {0}, z[6] = {0}, t0[6], t1[6], tau[6], r[6];
        for (int c = 0; c < C; c++) { e[c] = s->eps0[(long)c * N + v]; z[c] = e[c] + s->epsavg[c]; }  /* :993-999 */
        iso_mul(s, s->lamx[v], s->mux[v], e, t0);                         /* IsotropicMul_RF :1001 */
        iso_mul(s, lambda - s->lamx[v], mu - s->mux[v], z, t1);           /* IsotropicMul_RF_D :1002 */
        for (int c = 0; c < C; c++) tau[c] = t0[c] + t1[c];
        iso_inv(s, lambda, mu, tau, r);                                   /* IsotropicInv :1009 */
        for (int c = 0; c < C; c++) { s->tau0[(long)c * N + v] = t0[c]; s->estar[(long)c * N + v] = r[c]; }
    }
    double ne = 0.0;
    for (int c = 0; c < C; c++) ne += s->epsavg[c] * s->epsavg[c];
    const double normepsavg = sqrt(ne);                                   /* :1012 */
    int step = 0;
    for (;;) {
        step++;
        memcpy(s->epsolast, s->estar, sizeof(double) * (size_t)C * (size_t)N);   /* :1017-1022 */
        hom_solve(s);                                                     /* epso = total strain :1025 */
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) {
            double z[6] = {0}, t1[6], tau[6], r[6];
            for (int c = 0; c < C; c++) z[c] = s->epso[(long)c * N + v];
            iso_mul(s, lambda - s->lamx[v], mu - s->mux[v], z, t1);       /* :1027 */
            for (int c = 0; c < C; c++) tau[c] = s->tau0[(long)c * N + v] + t1[c];   /* :1030-1035 */
            iso_inv(s, lambda, mu, tau, r);                               /* :1038 */
            const double mix = 0.5;
            for (int c = 0; c < C; c++)
                s->estar[(long)c * N + v] = mix * r[c] + (1 - mix) * s->epsolast[(long)c * N + v];   /* :1040-1046 */
        }
        double avg = 0.0;                                                 /* AvgChange :957-975 */
        for (int c = 0; c < C; c++)
            for (long v = 0; v < N; v++) {
                const double d = s->estar[(long)c * N + v] - s->epsolast[(long)c * N + v];
                avg += d * d;
            }
        avg /= (double)N;
        const double convg = sqrt(avg);
        if ((convg < (1E-3) * normepsavg) || normepsavg < 1E-4) break;   /* :1049 */
        if (step > 100) break;                                            /* :1057 */
    }
    s->inhom_iters = step;
}

void ora_stress_solve(ora_state *s)
{
    if (s->lamx) stress_solve_inhom(s);
    else stress_solve_uniform(s);
}

void ora_set_moduli(ora_state *s, const double *lam, const double *mu)
{
    free(s->lamx); free(s->mux); free(s->estar);
    s->lamx = s->mux = s->estar = NULL;
    if (!lam || !mu) return;
    const size_t N = (size_t)s->N;
    s->lamx = (double *)malloc(N * sizeof(double));
    s->mux = (double *)malloc(N * sizeof(double));
    s->estar = (double *)malloc((size_t)s->C * N * sizeof(double));
    memcpy(s->lamx, lam, N * sizeof(double));
    memcpy(s->mux, mu, N * sizeof(double));
}

int ora_inhom_iterations(ora_state *s) { return s->inhom_iters; }

/* Print-step displacement: InhomogeneousStressSolver(..., uflag=2) (STZ/bmg.c:289-295)
   re-solves the stress and then runs HomogeneousStressSolver's displacement branch on the
   same round-tripped field (STZ/elastic.c:1050-1054, 873-900): per k,
   tmp_a = F_a + (-alpha)*(G*n_a) (HadamardSumnSProd, :765-777) and
   u_hat_a = (tmp_a.im*(1/mu)/|k|, -1.0*tmp_a.re*(1/mu)/|k|) (HadamardSiDivd, :754-763),
   u_a = Re(IFFT)/N + the supercell position (:893-899).  The reference adds the 2-D
   position with its slow index i along x and the fast index j along y
   (uo[0] += i*(1+e_xx)*LX + j*e_xy*LY); kept as written.  3-D: u_a += sum_b x_b
   (delta_ab + e_ab) L_b with x_b the grid index along axis b (x = the fast axis). */
void ora_displacement(ora_state *s, double *u)
{
    ora_stress_solve(s);
    const long N = s->N;
    const int C = s->C, D = s->p.dim;
    const int *n = s->p.n;
    const double lambda = s->lambda, mu = s->mu;
    const double alpha = (lambda + mu) / (lambda + 2.0 * mu);
    double *sp = s->spec;
    for (int c = 0; c < C; c++) {
        double *f = sp + 2 * (long)c * N;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) { f[2 * v] = s->epsolast[(long)c * N + v]; f[2 * v + 1] = 0.; }
        ora_fft3d(f, n[0], n[1], n[2], -1);
    }
    const double sdiv = 1.0 / mu;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        const int idx[3] = {(int)(v % n[0]), (int)((v / n[0]) % n[1]), (int)(v / ((long)n[0] * n[1]))};
        double kk[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) {
            double sa = (double)(idx[a]) * (double)(1.0 / n[a]);
            kk[a] = (sa - round(sa)) * ORA_TWO_PI;
        }
        double g2 = D == 2 ? sqrt(kk[0] * kk[0] + kk[1] * kk[1]) : sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
        if (idx[0] + idx[1] + idx[2] == 0) g2 = 1.0;
        const double g[3] = {s->g[0][v], s->g[1][v], s->g[2][v]};
        double tmp[2][3];
        for (int ri = 0; ri < 2; ri++) {
            double e[6] = {0}, t[6];
            for (int c = 0; c < C; c++) e[c] = sp[2 * ((long)c * N + v) + ri];
            iso_mul(s, lambda, mu, e, t);                 /* IsotropicMul_C :802 */
            double F[3], G;
            if (C == 3) {
                F[0] = t[0] * g[0] + t[2] * g[1];         /* :804 */
                F[1] = t[2] * g[0] + t[1] * g[1];         /* :805 */
                G = F[0] * g[0] + F[1] * g[1];            /* :806 */
            } else {
                F[0] = t[0] * g[0] + t[5] * g[1] + t[4] * g[2];
                F[1] = t[5] * g[0] + t[1] * g[1] + t[3] * g[2];
                F[2] = t[4] * g[0] + t[3] * g[1] + t[2] * g[2];
                G = F[0] * g[0] + F[1] * g[1] + F[2] * g[2];
            }
            for (int a = 0; a < D; a++) tmp[ri][a] = F[a] + -1.0 * alpha * (G * g[a]);
        }
        for (int a = 0; a < D; a++) {
            sp[2 * ((long)a * N + v)] = tmp[1][a] * sdiv / g2;
            sp[2 * ((long)a * N + v) + 1] = -1.0 * tmp[0][a] * sdiv / g2;
        }
    }
    const double L[3] = {s->p.lx, s->p.ly, s->p.lz};
    for (int a = 0; a < D; a++) {
        double *f = sp + 2 * (long)a * N;
        ora_fft3d(f, n[0], n[1], n[2], +1);
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) {
            const int idx[3] = {(int)(v % n[0]), (int)((v / n[0]) % n[1]), (int)(v / ((long)n[0] * n[1]))};
            const double *e = s->epsavg;
            double x = f[2 * v] / (double)N;
            if (D == 2) {
                const double i = (double)idx[1], j = (double)idx[0];
                if (a == 0) x += i * (1 + e[0]) * L[0] + j * e[2] * L[1];
                else x += i * e[2] * L[0] + j * (1 + e[1]) * L[1];
            } else {
                /* e_ab in Voigt order (xx, yy, zz, yz, xz, xy) */
                const int vo[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
                double pos = 0.0;
                for (int b = 0; b < 3; b++)
                    pos = pos + (double)idx[b] * ((a == b ? 1.0 : 0.0) + e[vo[a][b]]) * L[b];
                x += pos;
            }
            u[(long)a * N + v] = x;
        }
    }
}

/* ------------------------------------------------------------------ */
/* rates                                                                */
/* ------------------------------------------------------------------ */
void ora_softening(ora_state *s)
{
#pragma omp parallel for schedule(static) if(s->N >= ORA_PAR_MIN)
    for (long v = 0; v < s->N; v++)
        s->soft[v] = s->perm[v] + s->temp[v] * exp((s->tlast[v] - s->totime) / s->tTemp); /* :731 */
}

long long ora_act_energy(ora_state *s)
{
    const long N = s->N;
    const int C = s->C;
    long long nneg = 0;
    for (int m = 0; m < s->M; m++) {
#pragma omp parallel for schedule(static) reduction(+ : nneg) if(N >= ORA_PAR_MIN)
        for (long v = 0; v < N; v++) {
            double sg[6] = {0}, d[6] = {0}, pd[6] = {0};
            for (int c = 0; c < C; c++) sg[c] = s->sigma[(long)c * N + v];
            deps_of(s, m, v, d);
            deps_last_of(s, v, pd);
            double dwork = contract(s, sg, d) * s->voxel * 1e-18 * J_IN_EV;         /* :585-586 */
            double dot = contract(s, pd, d);                                       /* :613-614 */
            double norm1 = sqrt(contract(s, d, d));                                /* :615-616 */
            double norm2 = sqrt(contract(s, pd, pd));                              /* :617-618 */
            double dm = dot / (norm1 + ORA_EPSILON) / (norm2 + ORA_EPSILON);       /* :620 */
            dm = (1 + dm) * (1 + dm) / 4;                                          /* :621 */
            /* NO_SF_XTAL_MRO (:645-667): the modifier and the voxel's Softening shrink by
               0.1 / 0.01 / 0.9, Softening in place, so it compounds over the mode loop */
            const int t = s->mro[v];
            if (t) {
                const double f = t == 1 ? 0.1 : (t == 2 ? 0.01 : 0.9);
                dm *= f;
                s->soft[v] *= f;
            }
            double qv = s->q0[(long)m * N + v] * exp(-dm * s->soft[v]) - dwork / 2; /* :684 */
            s->q[(long)m * N + v] = qv;
            if (qv < 0) nneg++;                                                    /* bmg.c:152 */
        }
    }
    return nneg;
}

double ora_total_rate(ora_state *s)
{
    const long MN = (long)s->M * s->N;
    double t = 0;
#pragma omp parallel for schedule(static) if(MN >= ORA_PAR_MIN)
    for (long i = 0; i < MN; i++) s->rate[i] = s->ratepre * exp(-s->q[i] / s->kT);   /* STZ/evolution.c:125 */
    for (long i = 0; i < MN; i++) t += s->rate[i];                                  /* :126, sequential */
    return t;
}

/* normalise + cumulative + search on s->rate; STZ/evolution.c:150-224 */
int ora_select(ora_state *s, double dice, int *site, int *mode)
{
    const long MN = (long)s->M * s->N;
    double R = 0;
    for (long i = 0; i < MN; i++) R += s->rate[i];
    double tk = 1 / R;
#pragma omp parallel for schedule(static) if(MN >= ORA_PAR_MIN)
    for (long i = 0; i < MN; i++) s->rate[i] *= tk;
    for (long i = 1; i < MN; i++) s->rate[i] += s->rate[i - 1];
    long ev;
    if (s->rate[MN - 1] >= dice && s->rate[0] < dice) {
        long lo = 0, hi = MN - 1;           /* first i with S_i >= dice */
        while (lo < hi) {
            long mid = lo + (hi - lo) / 2;
            if (s->rate[mid] >= dice) hi = mid; else lo = mid + 1;
        }
        ev = lo;
    } else if (dice > 0 && dice <= s->rate[0]) {
        ev = 0;
    } else {
        return -1;   /* "KMC failed to locate a activation mode" */
    }
    *mode = (int)(ev / s->N);
    *site = (int)(ev - (long)(*mode) * s->N);
    return 0;
}

/* VoxelShearTransform glass branch, STZ/elastic.c:361-520 */
void ora_transform(ora_state *s, int mm, int site)
{
    const long N = s->N, v = site;
    double d[6];
    deps_of(s, mm, v, d);
    for (int k = 0; k < s->S; k++) s->deps_last[(long)k * N + v] = s->deps[((long)mm * s->S + k) * N + v];
    for (int c = 0; c < s->C; c++) s->eps0[(long)c * N + v] += d[c];            /* :377-379 */
    {
        double sg[6] = {0};
        for (int c = 0; c < s->C; c++) sg[c] = s->sigma[(long)c * N + v];
        const double dh = contract(s, sg, d) * 1E9;                              /* :381-382, J/m^3 */
        s->tlocal[v] += fabs(dh) * ORA_THERMAL_COEFF1;                          /* :383 */
    }
    double ds = contract(s, d, d);                                             /* :385-386 */
    const int six = s->mro[v] == 1;                                            /* PermSoftmro, TempSoftmro */
    s->perm[v] += ds * (six ? 20.0 : s->permsoft);                             /* :387-390; STZ/init.c:60-63 */
    if (s->perm[v] > s->permcap) s->perm[v] = s->permcap;                      /* :391 */
    s->temp[v] = ds * (six ? 60.0 : s->tempsoft);                              /* :393-396 */
    s->tlast[v] = s->totime;                                                   /* :397 */
    if (s->mro[v] == 0) {
        for (int m = 0; m < s->M; m++) draw_deps(s, m, v);                     /* :508-512 */
    } else {
        /* crystal-like MRO: new barriers, the strains stay (:446-506) */
        for (int m = 0; m < s->M; m++)
            s->q0[(long)m * N + v] = s->p.mro_q0 * (1 - s->p.random_ratio + s->p.random_ratio * (ora_mt_uniform(&s->rng)));
    }
}

double ora_elastic_energy(ora_state *s)
{
    const long N = s->N;
    double e = 0.;
    for (long v = 0; v < N; v++) {
        double sg[6] = {0}, el[6];
        for (int c = 0; c < s->C; c++) sg[c] = s->sigma[(long)c * N + v];
        isotropicinv(s, sg, el);
        e += 0.5 * contract(s, sg, el);   /* STZ/elastic.c:544 */
    }
    return e;
}

/* athermal error subtraction, STZ/evolution.c:91-97 (reference) or loading-consistent */
static void athermal_correction(ora_state *s)
{
    double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
    if (s->p.athermal_correction == 0) {
        if (s->C == 3) { ts[0] = 0.0; ts[1] = s->stress[1]; ts[2] = s->stress[2]; }
        else { ts[0] = 0.0; for (int c = 1; c < 6; c++) ts[c] = s->stress[c]; }
    } else {
        for (int c = 0; c < s->C; c++) ts[c] = (c == s->lc) ? 0.0 : s->stress[c];
    }
    isotropicinv(s, ts, te);
    for (int c = 0; c < s->C; c++) s->epsavg[c] -= te[c];
}

/* HeatConducting(dt), STZ/elastic.c:549-582: Tlocal through the full complex transform,
   both parts times exp(-g2*dt) (g2 = |k| of the reference grid, 1 at k = 0), back, /N.
   Returns -4 on a NaN, as the reference aborts with -4. */
int ora_heat_conduct(ora_state *s, double dt)
{
    const long N = s->N;
    double *f = s->spec;                      /* scratch: one complex field */
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) { f[2 * v] = s->tlocal[v]; f[2 * v + 1] = 0.0; }
    ora_fft3d(f, s->p.n[0], s->p.n[1], s->p.n[2], -1);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(|:bad) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        const double m = exp(-s->g2[v] * dt);
        f[2 * v] *= m;
        f[2 * v + 1] *= m;
        if (isnan(f[2 * v]) || isnan(f[2 * v + 1])) bad = 1;
    }
    if (bad) return -4;
    ora_fft3d(f, s->p.n[0], s->p.n[1], s->p.n[2], +1);
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) s->tlocal[v] = f[2 * v] / (double)N;
    return 0;
}

void ora_set_heat(ora_state *s, int on) { s->heat = on != 0; }

static int athermal(ora_state *s, ora_step_info *info)
{
    const long N = s->N;
    int ntr = 0;
    s->astep++;
    int *pick = (int *)malloc(sizeof(int) * (size_t)N);
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
    for (long v = 0; v < N; v++) {
        int mm = 0;
        double minq = s->q[v];
        for (int m = 0; m < s->M; m++)
            if (minq > s->q[(long)m * N + v]) { mm = m; minq = s->q[(long)m * N + v]; }
        pick[v] = (minq < 0 || fabs(minq) < s->p.athermal_threshold) ? mm : -1;
    }
    for (long v = 0; v < N; v++) {
        const int mm = pick[v];
        if (mm >= 0) {
            double sg[6] = {0}, d[6] = {0};
            for (int c = 0; c < s->C; c++) sg[c] = s->sigma[(long)c * N + v];
            deps_of(s, mm, v, d);
            double dwork = contract(s, sg, d) * s->voxel * 1e-18 * J_IN_EV;
            s->soc_elast += fabs(dwork);
            ora_transform(s, mm, (int)v);
            s->avlan_counter++;
            ntr++;
        }
    }
    free(pick);
    ora_stress_solve(s);
    athermal_correction(s);
    info->kind = ORA_ATHERMAL;
    info->n_transformed = ntr;
    return 0;
}

static int kmc(ora_state *s, ora_step_info *info)
{
    const long N = s->N;
    const int C = s->C, lc = s->lc;
    double R = ora_total_rate(s);
    double t_kmc = 1 / R;
    info->total_rate = R;
    info->dt_kmc = t_kmc;
    const double sr = s->p.strainrate;
    if (t_kmc < s->dt) {
        s->kstep++;
        double dice = ora_mt_uniform(&s->rng);
        int site, mode;
        if (ora_select(s, dice, &site, &mode)) return -1;
        double qch = s->q[(long)mode * N + site];
        ora_transform(s, mode, site);
        if (s->heat && ora_heat_conduct(s, ORA_THERMAL_DIFFUSIVITY * t_kmc)) return -4;   /* STZ/evolution.c:247-249 */
        ora_stress_solve(s);
        double ts[6] = {0, 0, 0, 0, 0, 0}, te[6];
        if (s->p.loading == 1) {              /* SHEAR, STZ/evolution.c:298-315 */
            for (int c = 0; c < C; c++) ts[c] = (c == lc) ? 0 : s->stress[c];
            isotropicinv(s, ts, te);
            for (int c = 0; c < C; c++)
                if (c != lc) s->epsavg[c] += -te[c];
            s->epsavg[lc] += sr * t_kmc;
            s->stress[lc] += 2 * s->mu * sr * t_kmc;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
            for (long v = 0; v < N; v++) s->sigma[(long)lc * N + v] += 2 * s->mu * sr * t_kmc;
        } else {                              /* TENSION, STZ/evolution.c:281-297 */
            ts[0] = 0.0;
            for (int c = 1; c < C; c++) ts[c] = s->stress[c];
            isotropicinv(s, ts, te);
            s->epsavg[0] += (sr * t_kmc * 1 - te[0]);
            s->epsavg[1] += (sr * t_kmc * (-s->nu) - te[1]);
            if (C == 3) s->epsavg[2] += (-te[2]);
            else {
                s->epsavg[2] += (sr * t_kmc * (-s->nu) - te[2]);
                for (int c = 3; c < 6; c++) s->epsavg[c] += (-te[c]);
            }
            s->stress[0] += s->E * sr * t_kmc;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
            for (long v = 0; v < N; v++) s->sigma[v] += s->E * sr * t_kmc;
        }
        s->totime += t_kmc;
        info->kind = ORA_THERMAL;
        info->site = site;
        info->mode = mode;
        info->q_chosen = qch;
        info->mro = s->mro[site];                                              /* chosenM, :206, 220 */
    } else {
        s->estep++;
        s->totime += s->dt;
        if (s->heat && ora_heat_conduct(s, ORA_THERMAL_DIFFUSIVITY * s->dt)) return -4;  /* STZ/evolution.c:325-327 */
        const double inc = s->p.strainincrement;
        if (s->p.loading == 1) {              /* STZ/evolution.c:349-356 */
            s->epsavg[lc] += sr * 1;
            s->stress[lc] += 2 * s->mu * inc;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
            for (long v = 0; v < N; v++) s->sigma[(long)lc * N + v] += 2 * s->mu * inc;
        } else {                              /* STZ/evolution.c:335-347 */
            s->epsavg[0] += (inc * 1);
            s->epsavg[1] += (inc * (-s->nu));
            if (C == 3) s->epsavg[2] += (inc * 0.0);
            else { s->epsavg[2] += (inc * (-s->nu)); for (int c = 3; c < 6; c++) s->epsavg[c] += (inc * 0.0); }
            s->stress[0] += s->E * inc;
#pragma omp parallel for schedule(static) if(N >= ORA_PAR_MIN)
            for (long v = 0; v < N; v++) s->sigma[v] += s->E * inc;
        }
        info->kind = ORA_ELASTIC;
    }
    return 0;
}

int ora_step(ora_state *s, ora_step_info *info)
{
    memset(info, 0, sizeof(*info));
    s->step++;
    if (s->p.softening == 1) ora_softening(s);
    long long nneg = ora_act_energy(s);
    info->n_negative = nneg;
    int rc;
    if (nneg > 0) rc = athermal(s, info);
    else {
        /* UpdateSOC resets the avalanche counter (STZ/evolution.c:5-45) */
        s->avlan_counter = 0;
        s->soc_elast = 0;
        rc = kmc(s, info);
    }
    if (rc) return rc;
    int term = 0;
    if (s->print_on) {
        /* print block, STZ/bmg.c:181-380; with printperiod <= 0 the test holds at every step */
        double cur = s->epsavg[s->lc];
        if (fabs(cur - s->strainhist_last) >= (s->printperiod * s->p.strainincrement * 0.99)) {
            s->strainhist_last = cur;                                     /* :193-201 */
            info->el_energy = ora_elastic_energy(s);                      /* :174-178, recorded :225 */
            memcpy(s->q0, s->q1, sizeof(double) * (size_t)s->M * (size_t)s->N);   /* Q0 <- Q1, :215-223 */
            ora_stress_solve(s);              /* print-step re-solve, STZ/bmg.c:289-295 */
            info->printed = 1;
        }
        if (fabs(s->epsavg[s->lc]) > s->endstrain) term = 1;
        if (s->p.loading == 0 && s->stress[0] < -0.5) term = 1;
        if (s->p.loading == 1 && fabs(s->stress[s->lc]) < ORA_EPSILON) term = 1;
    }
    info->totime = s->totime;
    for (int c = 0; c < 6; c++) { info->epsavg[c] = s->epsavg[c]; info->stress[c] = s->stress[c]; }
    return term;
}
