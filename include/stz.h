/*
 * stz.h -- C-ABI of the MI355X-native STZ kinetic Monte Carlo step.
 *
 * The reference simulator ("Mesoscale simulation STZ model/", abbreviated STZ/)
 * has no plugin API: its hot path sits behind plain C prototypes acting on
 * module globals (STZ/bmg.h:149-152, 295-311) and aborts the MPI job on error.
 * This header replaces those prototypes with an explicit context, plain
 * pointers and sizes, and negative return codes instead of MPI_Abort.
 *
 * Conventions
 *   - Every function returns STZ_OK (0) or a negative code.
 *   - Host buffers are caller-owned and read/written only during the call.
 *   - One context per host thread; a context is not internally thread-safe.
 *   - Doubles cross the ABI in both precisions.
 *   - Field layouts (N = n[0]*n[1]*n[2] voxels, voxel index
 *     v = (i2*n[1] + i1)*n[0] + i0, i0 fastest; C = 3 (2-D) or 6 (3-D)
 *     tensor components xx, yy, xy  or  xx, yy, zz, yz, xz, xy;
 *     S = 2 or 5 stored transformation-strain components
 *     xx, xy  or  xx, yy, yz, xz, xy; M = modes):
 *       EPS0 [C][N], SIGMA [C][N], Q0 [M][N], DEPS [M][S][N],
 *       DEPS_LAST [S][N], PERM/TEMP/TLAST [N].
 */
#ifndef STZ_H
#define STZ_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes; -1 and -4 mirror the reference's MPI_Abort codes */
#define STZ_OK 0
#define STZ_E_SELECT (-1)      /* KMC failed to locate an activation mode (STZ/evolution.c:186-234) */
#define STZ_E_HIP (-2)         /* HIP runtime error */
#define STZ_E_ARG (-3)         /* bad argument / unsupported configuration */
#define STZ_E_NAN (-4)         /* NaN in the stress solve (STZ/elastic.c:212-215) */
#define STZ_E_UNSUPPORTED (-5) /* crystal-like MRO voxel drawn, or an unsupported combination */
#define STZ_E_TOO_LARGE (-6)   /* stz_rates_noalloc: modes * ceil(voxels / 256) above 2^21;
                                  stz_persist_probe: a table beyond its 31-bit counters */

#define STZ_FP64 0
#define STZ_MIXED 1
#define STZ_TENSION 0
#define STZ_SHEAR 1
#define STZ_SOLVER_FFT 0     /* full Green's-function solve after every transform */
#define STZ_SOLVER_LOCAL 1   /* event-local update, full solve every `resync` updates */

/* fields for stz_upload / stz_download */
#define STZ_F_EPS0 0
#define STZ_F_SIGMA 1      /* download applies the pending uniform stress increment */
#define STZ_F_Q0 2
#define STZ_F_DEPS 3
#define STZ_F_DEPS_LAST 4
#define STZ_F_PERM 5
#define STZ_F_TEMP 6
#define STZ_F_TLAST 7
#define STZ_F_MINQ 8       /* download only: per-voxel min over modes of Q from the last stz_rates */
#define STZ_F_TLOCAL 9     /* local temperature Tlocal (STZ/init.c:470-476), FP64 */
#define STZ_F_MRO 10       /* download only: MRO type per voxel (0 glass, 1 six-, 2 two-, 3 ten-fold) */

/* step kinds */
#define STZ_STEP_ELASTIC 0
#define STZ_STEP_THERMAL 1
#define STZ_STEP_ATHERMAL 2

/*
 * Run parameters.  The reference fixes most of these at compile time
 * (STZ/bmg.h:22-98, STZ/bmgFlags.h) or in the 17-field input file
 * (STZ/io.c:4-101); here they are runtime values.
 */
typedef struct stz_params {
    int dim;                 /* 2 (the reference) or 3 */
    int n[3];                /* grid; n[0] fastest (reference NY), n[1] (reference NX), n[2] (1 in 2-D) */
    int modes;               /* MODES, STZ/bmg.h:29 */
    int precision;           /* STZ_FP64 | STZ_MIXED */
    int loading;             /* STZ_TENSION (bmgFlags TENSION) | STZ_SHEAR (bmgFlags SHEAR) */
    int athermal_correction; /* 0: subtract C0^-1(0, <s_yy>, <s_xy>) as STZ/evolution.c:91-97;
                                1: subtract every mean-stress component except the loaded one */
    int softening;           /* "Softening Flag", STZ/io.c:53-56 */
    int seed;                /* "Random seed", STZ/io.c:33-36 */
    double E, nu;            /* "Moduli(E, nu)", STZ/io.c:7-12 */
    int double_nu;           /* 1: LOW_T doubling of nu, STZ/init.c:26 */
    double q0_mean;          /* "Stress-free energy barrier", STZ/io.c:14-18 */
    double q0_std;           /* 0.25 hard-coded at STZ/init.c:185 */
    double gamma0;           /* GAMMA0, STZ/bmg.h:94 */
    double temperature;      /* T__K, STZ/bmg.h:36 */
    double trial_frequency;  /* TRIAL_FREQUENCY, STZ/bmg.h:42 */
    double lx, ly, lz;       /* voxel edges (nm), STZ/bmg.h:84-90 */
    double strainrate;       /* "Applied strain rate" */
    double strainincrement;  /* "Strain increment" */
    double perm_soft, temp_soft, perm_cap; /* STZ/init.c:60-68 */
    double temp_soft_barrier;              /* 0.37 eV in tTemp, STZ/init.c:49 */
    double athermal_threshold;             /* 0.33 eV, STZ/evolution.c:65 */
    /* crystal-like medium-range order (2-D only; all 0: glass everywhere) */
    double mro_fraction;     /* "Volume fraction of crystal-like MRO", STZ/io.c:83-86 (InitMRO 0) */
    double mro_q0;           /* "Stress-free energy barrier for crystal-like MRO", STZ/io.c:93-97 */
    double random_ratio;     /* "Random ratio of fluctuation in energy barrier", STZ/io.c:37-40 */
} stz_params;

typedef struct stz_step_result {
    int kind;                /* STZ_STEP_* */
    int site, mode;          /* thermal event */
    int n_transformed;       /* athermal: voxels transformed */
    long long n_negative;    /* #(mode, voxel) with Q < 0 (STZ/bmg.c:147-157) */
    double q_chosen;         /* Q of the selected (mode, site), written to actQ.out */
    double total_rate;       /* R */
    double dt_kmc;           /* 1/R */
    double totime;           /* clock after the step */
    double epsavg[6];        /* EpsAvg after the step */
    double stress[6];        /* Stress (<sigma>) after the step */
    int printed;             /* the print block ran and re-solved the stress */
    double soc_elast;        /* athermal: sum |DWork| of the transformed voxels (STZ/evolution.c:66-68) */
} stz_step_result;

typedef struct stz_ctx stz_ctx;

/* Allocates device state for p (fields zero, tlast = -100*tTemp) on the current HIP device.
   Replaces Init + FFT_MPI_Init + Grad + Alloc (STZ/init.c:6-140, 225-479). */
int stz_create(const stz_params *p, stz_ctx **out);
void stz_destroy(stz_ctx *ctx);
const char *stz_last_error(const stz_ctx *ctx);

/* Draws the initial state from the context's MT19937 stream in the reference order
   (MRO map, Energy_Init, DStrain: STZ/init.c:350-397, 181-198; STZ/elastic.c:60-138)
   and runs the initial stress solve (STZ/bmg.c:117-127).  The MRO map is InitMRO 0: a
   voxel whose uniform is <= mro_fraction is six-fold crystal-like MRO, oriented by a
   second uniform (2-D only). */
int stz_init_state(stz_ctx *ctx);

/* InitMRO 1 (STZ/init.c:398-431): stz_init_state with the MRO map read from
   MROstructure.ms by the caller -- type[v] in {0 glass, 1 six-fold, 2 two-fold, 3 ten-fold}
   and the orientation MROthe0 theta[v] per voxel in the reference's voxel order; ten-fold
   voxels turn two-fold with probability 0.5 (one uniform each, voxel order), then Q0, the
   transformation strains of each family (STZ/elastic.c:60-138) and the initial solve.
   Crystal-like MRO is 2-D only (STZ_E_UNSUPPORTED in 3-D).  With MRO voxels, steps run on
   the host path (stz_run_steps falls back to it) and ensembles return STZ_E_UNSUPPORTED;
   a crystal-like voxel's transform redraws its M barriers from M uniforms (the first M of
   the M*S values stz_apply_event / stz_athermal_apply take per voxel). */
int stz_init_state_mro(stz_ctx *ctx, const int *type, const double *theta);

/* Initial state drawn on the GPU for large grids, where the reference's sequential
   host stream (stz_init_state) would take minutes: the same distributions from the
   counter-based Philox4x32-10 generator keyed by seed, with every value a function of
   (seed, voxel, mode, component), so the FP64 and mixed runs of one seed start from
   the same state.  Voxels within one of nsph spheres (spheres[4*i..]: centre x, y, z in
   voxels and radius, periodic) draw Q0 ~ N(inc_mean, inc_std) -- a soft inclusion;
   *inside (optional) receives their count.  Then epsavg = 0 and the initial solve.
   No counterpart in the reference. */
int stz_init_state_device(stz_ctx *ctx, unsigned long long seed, const double *spheres, int nsph, double inc_mean,
                          double inc_std, long long *inside);

/* Field transfer (host buffers of doubles, read or written only during the call).  An
   upload of STZ_F_Q0, _DEPS, _DEPS_LAST, _PERM, _TEMP or _TLAST discards the rate pass's cached
   Q0*exp(-D*S) values, which are recomputed by the next rate pass.  An upload of STZ_F_Q0
   also sets the stress-free barriers Q1 that every print step restores (STZ/init.c:189,
   STZ/bmg.c:215-223). */
int stz_upload(stz_ctx *ctx, int field, const double *host);
int stz_download(stz_ctx *ctx, int field, double *host);
/* The host scalars EpsAvg, Stress (epsavg/stress may be NULL to keep them) and totime.
   A new totime also discards the cached Q0*exp(-D*S) values: the rate pass keeps a
   voxel's cached softening without re-reading Perm/Temp/tlast once its temporary term
   is zero, which holds only while the clock advances. */
int stz_set_scalars(stz_ctx *ctx, const double epsavg[6], const double stress[6], double totime);
int stz_get_scalars(stz_ctx *ctx, double epsavg[6], double stress[6], double *totime);

/* InhomogeneousStressSolver(..., uflag=0) (STZ/elastic.c:977-1066, 779-828): sigma field
   and <sigma> from EPS0 and EpsAvg.  With uniform moduli one homogeneous solve; with
   per-voxel moduli (stz_set_moduli) the reference's fixed-point loop of homogeneous solves
   (mix 0.5, until the RMS change of the reference-medium eigenstrain is below
   1e-3 |EpsAvg| or |EpsAvg| < 1e-4, at most 101 solves).  Discards a pending uniform
   stress increment (the reference's solve overwrites Sigma). */
int stz_stress_solve(stz_ctx *ctx);

/* Per-voxel Lame moduli Lambda(x), Mu(x) (N doubles each, voxel order) of an elastically
   inhomogeneous glass: the Lambda/Mu arguments of every InhomogeneousStressSolver call
   (STZ/bmg.c:93, 121, 292; STZ/evolution.c:83, 253), set in the reference by
   STZ/init.c:457-465 (there uniform: Lambda = lambda, Mu = mu).  lambda = mu = NULL
   returns to uniform moduli.  The reference medium C0 stays (lambda, mu) of the
   parameters.  Every later stress solve runs the fixed-point loop; steps run on the host
   path (stz_step; stz_run_steps falls back to it); the event-local update and ensembles
   return STZ_E_UNSUPPORTED.  Device memory: (2 + 3C) doubles per voxel. */
int stz_set_moduli(stz_ctx *ctx, const double *lambda, const double *mu);

/* HEAT_CONDUCT (STZ/bmgFlags.h:84; off in the reference's build): with on = 1, every
   thermal step (after its transform) and every elastic step diffuses the local
   temperature STZ_F_TLOCAL by HeatConducting(thermal_diffusivity * dt)
   (STZ/evolution.c:247-249, 325-327; STZ/elastic.c:549-582).  Tlocal is heated by every
   transform whether or not the flag is on (STZ/elastic.c:381-383) and feeds nothing else.
   With heat conduction steps run on the host path (stz_run_steps falls back to it) and
   ensembles return STZ_E_UNSUPPORTED.  No counterpart as a call in the reference. */
int stz_set_heat_conduct(stz_ctx *ctx, int on);
/* One HeatConducting(dt) of STZ_F_TLOCAL: spectrum times exp(-|k| dt), |k| of the
   reference's k grid and 1 at k = 0.  STZ_E_NAN on a NaN (the reference aborts with -4). */
int stz_heat_conduct(stz_ctx *ctx, double dt);

/* Homogeneous solves the last stress solve took (1 with uniform moduli).  No counterpart
   in the reference (its loop counter `step`, STZ/elastic.c:1014-1015). */
int stz_inhom_iterations(stz_ctx *ctx);

/* Event-local form of the same solve, replacing InhomogeneousStressSolver(uflag=0) after
   transforms at `n` distinct sites (STZ/evolution.c:70-89, 250-259): eps0 has already changed by
   the transformation strains that the transforms recorded (the PreDEps0 snapshot of
   STZ/elastic.c:369-374) at those sites and nowhere else.  The solve is linear and
   translation invariant, so the new stress is the field as last solved, plus C0 times the
   change of EpsAvg since then, plus the superposed impulse responses of the sites (built
   once per grid and elastic constants by FP64 FFT solves of unit impulses and kept for
   the half space of offsets, since the reference's operator is even: about S*C*N/2
   values in the storage precision, shared by contexts of one grid and elasticity).  Updates the sigma field and <sigma> like stz_stress_solve, to
   rounding.  No counterpart in the reference.  n <= 64; -3 if EPS0 or SIGMA was uploaded
   since the last full solve. */
int stz_stress_update_local(stz_ctx *ctx, int n, const int *sites);

/* Print-step displacement field (InhomogeneousStressSolver(..., uflag=2) at STZ/bmg.c:289-295:
   HomogeneousStressSolver's displacement branch, STZ/elastic.c:873-900): from the
   round-tripped EPS0 and EpsAvg, u_hat = -i (F - alpha G n) / (mu |k|) and
   u = Re(IFFT)/N + the supercell position (2-D as the reference writes it, its slow
   index along x; 3-D: x along the fast axis).  host_u receives dim*N doubles
   [component][voxel] (the X1R/X2R outputs, STZ/bmg.c:300-330).  The stress is not
   changed, except with per-voxel moduli: then the call first re-solves (the sigma field
   and <sigma>), as the reference's uflag=2 call does, and u comes from the eigenstrain
   the fixed point ends on (its epsolast, STZ/elastic.c:1051-1055). */
int stz_displacement(stz_ctx *ctx, double *host_u);

/* How the resident step (stz_run_steps, stz_ensemble_run) brings the stress up to date
   after transforms: STZ_SOLVER_FFT (the full solve on the device after every transform),
   or STZ_SOLVER_LOCAL: the event-local update for a thermal event and for avalanches of
   at most local_max voxels; a step that needs the full solve instead (an avalanche
   larger than local_max, or `resync` local updates since the last full solve) halts
   before changing anything and runs on the host path (stz_step), as do print steps.
   The choice depends only on the run's own state, so results do not depend on batching.
   No counterpart in the reference.  Defaults: FFT, 1024, 16. */
int stz_set_solver(stz_ctx *ctx, int mode, long long resync, int local_max);

/* SofteningFactor + ActEnergy + Q<0 count + KMC rate total
   (STZ/init.c:711-739, 561-690; STZ/bmg.c:147-157; STZ/evolution.c:120-134). */
int stz_rates(stz_ctx *ctx, long long *n_negative, double *total_rate);

/* KMC normalise + prefix + search (STZ/evolution.c:136-224) on the rates of the last
   stz_rates call: first event index i = mode*N + site whose cumulative rate reaches
   dice*R.  q receives Q(mode, site). */
int stz_select(stz_ctx *ctx, double dice, int *site, int *mode, double *q);

/* VoxelShearTransform glass branch (STZ/elastic.c:361-520) at (site, mode) with the new
   transformation strains of all M modes of the voxel (deps_new[m*S + s], already scaled
   by 0.5*gamma0).  Uses the context clock for tlast. */
int stz_apply_event(stz_ctx *ctx, int site, int mode, const double *deps_new);

/* athermal_plasticity scan (STZ/evolution.c:52-77): number of voxels whose minimum Q
   over modes is < 0 or within the threshold, in row-major order. */
int stz_athermal_scan(stz_ctx *ctx, int *n);
int stz_athermal_list(stz_ctx *ctx, int *sites, int *modes, int cap);
/* applies the scanned transforms; deps_new holds n*M*S values, voxel by voxel in list order.
   soc_elast receives sum |DWork| of the transforms (STZ/evolution.c:66-68). */
int stz_athermal_apply(stz_ctx *ctx, const double *deps_new, double *soc_elast);

/* Sigma[comp][*] += delta (STZ/evolution.c:293-296, 311-314, 344-356), applied lazily
   before the next rate evaluation. */
int stz_add_uniform_stress(stz_ctx *ctx, int comp, double delta);

/* ElasticEnergy sum over voxels (STZ/elastic.c:533-547). */
int stz_elastic_energy(stz_ctx *ctx, double *e);

/* One iteration of the reference main loop (STZ/bmg.c:138-411) driven by the context's
   own MT19937 stream: softening, ActEnergy, athermal or thermal/elastic KMC, loading,
   print-period re-solve.  Returns 1 when the run terminates (end strain / instability).
   The print block and the termination tests run after every step once stz_set_print has
   been called (a context without it never prints or terminates); as in the reference, a
   printperiod <= 0 makes every step a print step (STZ/bmg.c:181-187). */
int stz_set_print(stz_ctx *ctx, int printperiod, double endstrain);
int stz_step(stz_ctx *ctx, stz_step_result *res);

/* stz_step in two halves, for callers that write the print-block outputs from the state
   before the print-step re-solve as the reference does (STZ/bmg.c:181-287):
   stz_step_kmc runs the iteration up to the print block and sets res->printed = 1 when
   the print block is due (it then returns 0 and the caller must call stz_step_print);
   stz_step_print resets the barriers Q0 <- Q1 (STZ/bmg.c:215-223: only crystal-like
   transforms change Q0), re-solves the stress (STZ/bmg.c:289-295) and checks termination. */
int stz_step_kmc(stz_ctx *ctx, stz_step_result *res);
int stz_step_print(stz_ctx *ctx, stz_step_result *res);

/* Runs up to nsteps iterations of the reference's while(1) loop body (STZ/bmg.c:138-411)
   device-resident: rates, step kind, selection, transforms, stress solve, loading, clock
   and the print/termination tests run on the GPU without a host round trip, consuming
   the context's MT19937 stream by index, so the results equal those of nsteps calls of
   stz_step.  Print steps are re-solved on the host as in stz_step.  Stops early on
   termination (return 1) or error (< 0).  results (optional) receives one entry per
   completed step; done receives the number of completed steps.  nsteps = 0 only
   prepares the resident resources and the random-number lookup window. */
int stz_run_steps(stz_ctx *ctx, int nsteps, stz_step_result *results, int *done);

/* Resident-run settings: steps per host synchronisation (batch, 0 = keep; default 64),
   use_graph = g > 0: replay g consecutive steps per hipGraph launch, 0: direct launches
   (-1 = keep; default 0), words of the MT19937 lookup window (0 = keep; default 2^21).
   No counterpart in the reference. */
int stz_run_config(stz_ctx *ctx, int batch, int use_graph, long long window_words);
/* With on = 1, stz_run_steps returns right after a print step (last res.printed = 1) and
   before its re-solve, like stz_step_kmc: the caller writes the print block's outputs
   (STZ/bmg.c:181-287) and then calls stz_step_print. */
int stz_run_stop_at_print(stz_ctx *ctx, int on);

/* Batched ensemble: replicas (contexts with their own grid, seed, temperature and barrier
   spread) advanced together, every kernel of a step launched once per shape class over
   all its replicas (directly or replayed from a hipGraph, as stz_run_config).  The contexts stay owned by the
   caller and must outlive the ensemble.  stz_ensemble_run advances each replica by up to
   nsteps steps exactly as stz_run_steps would: results[i] (optional, nsteps entries)
   receives replica i's records, done[i] their count and rcs[i] its return code.  No
   counterpart in the reference (one run per process there). */
typedef struct stz_ensemble stz_ensemble;
int stz_ensemble_create(stz_ctx *const *members, int n, stz_ensemble **out);
int stz_ensemble_config(stz_ensemble *e, int batch, int use_graph);
int stz_ensemble_run(stz_ensemble *e, int nsteps, stz_step_result *const *results, int *done, int *rcs);
void stz_ensemble_destroy(stz_ensemble *e);

/* Persistent step (on = 1): each step of every replica of every shape class is ONE kernel
   launch instead of a sequence of launches per class.  Its workgroups take tickets of the
   step's work in dependency order (rates, group sums, chunk totals, decision, athermal
   pass, event-local update, bookkeeping) and wait for their replica's earlier phases on
   device counters; replayed from a hipGraph with use_graph as the kernel sequence is.
   Results equal those of the kernel sequence bit for bit.  Takes replicas of one precision
   and dimension on the event-local update (stz_set_solver STZ_SOLVER_LOCAL), at most 8
   shape classes: STZ_E_UNSUPPORTED otherwise (a step that needs the full solve still halts
   to the host path, as without it).  on = 0 returns to the kernel sequence.  No
   counterpart in the reference (its step is a sequence of MPI-parallel function calls,
   STZ/bmg.c:138-411). */
int stz_ensemble_persistent(stz_ensemble *e, int on);
int stz_run_persistent(stz_ctx *ctx, int on);
/* CPU check of the persistent step's ticket table (no device call; for tests): builds the
   table of ncls shape classes in launch order, shape[7*c ..] = {dim, n0, n1, n2, modes,
   precision, replicas}, decodes every ticket as the kernel does and counts violations (an
   index out of its range, an item covered other than exactly once, a ticket preceding a
   ticket of an earlier phase of its replica).  out = {tickets, items, violations};
   STZ_E_TOO_LARGE when the table does not fit the 31-bit counters. */
int stz_persist_probe(int ncls, const int *shape, long long out[3]);

/* Autotuner of the stress-solve schedule (no counterpart in the reference): times every
   compiled variant of the four solve kernel slots -- x-pass rows (forward, inverse),
   y-pass columns, fused z-pass columns, each with the length's radix plans -- with HIP
   events on the context's own fields (the stress field is saved and restored), keeps
   the fastest and caches the choice per grid shape and precision (flags & 1: measure
   again).  best_ms (optional) receives the chosen variants' mean times (-1 when taken
   from the cache).  get/set read or pin the choices (0 = built-in default).  A choice is
   tile | plan << 8: the tile (rows or columns) in the low byte, the radix plan (0 = the
   length's default factorisation, 1..3 alternatives such as 128 = 16*8 or
   256 = 4*8*8) above it. */
#define STZ_TUNE_MAXCAND 16
int stz_autotune(stz_ctx *ctx, int reps, int flags, double best_ms[4]);
int stz_get_tuning(stz_ctx *ctx, int out[4]);
int stz_set_tuning(stz_ctx *ctx, const int in[4]);
/* compiled variants of one slot (0 x forward rows, 1 y columns, 2 fused columns, 3 x
   inverse rows) for this context's shape; returns their count */
int stz_tune_candidates(stz_ctx *ctx, int slot, int out[STZ_TUNE_MAXCAND]);

/* Host evaluation of the Philox4x32-10 rounds used by stz_init_state_device (known-answer
   checks; no GPU needed). */
int stz_philox_probe(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);

/* Host-only check of the MT19937 lookup window used by the resident step: window values
   and stream positions against the same draws taken in order (no GPU needed). */
int stz_mt_window_probe(unsigned long seed, long long skip, long long window_words, long long offset, int with_uniform,
                        int ngauss, double *win_out, double *seq_out, long long *win_end, long long *seq_end);

/* Host RNG of the context (GSL mt19937 + gsl_ran_gaussian polar method). */
int stz_rng_seed(stz_ctx *ctx, unsigned long seed);
int stz_rng_uniform(stz_ctx *ctx, double *u);
int stz_rng_gaussian(stz_ctx *ctx, double sigma, double *g);

/* Checkpoint and restart (no counterpart in the reference, which never saves its RNG
   state, Q0, transformation strains or softening state).  The file holds the MT19937
   state and position, every per-voxel array in its storage precision, EpsAvg, <sigma>,
   the clock, the pending uniform stress, the print/termination state and the counters;
   loading it into a context created with the same grid, mode count and precision
   continues the run exactly as the saving context would (-3 on a mismatch or a corrupt
   file; after a failed load the context's state is undefined until the next successful
   load or initialisation). */
int stz_checkpoint_save(stz_ctx *ctx, const char *path);
int stz_checkpoint_load(stz_ctx *ctx, const char *path);

/* Counters: step, astep, kstep, estep, transforms, solves. */
int stz_counters(stz_ctx *ctx, long long out[6]);

/* Plain device copy of `bytes` (rounded down to 16-byte words) on the current device,
   timed with HIP events over `reps` launches: avg_ms per launch and gbs = 2*bytes/time
   (read + write).  Measurement hook with no counterpart in the reference: the stream-copy
   bandwidth reported beside the kernels' rooflines, and the kernel whose byte counters
   calibrate the hardware-counter traffic figures.  Allocates and frees its buffers. */
int stz_copy_probe(long long bytes, int reps, double *avg_ms, double *gbs);

/* MT19937 state (624 words + position) of the context stream, for checkpoints and for
   handing a run over to another implementation mid-trajectory.  No counterpart in the
   reference, whose RNG state is never saved. */
int stz_rng_get_state(stz_ctx *ctx, unsigned int words[625]);
int stz_rng_set_state(stz_ctx *ctx, const unsigned int words[625]);

/* Run-control scalars: out[0] = last printed strain (strainhist[N_print-1], STZ/bmg.c:182),
   out[1] = print period, out[2] = end strain. */
int stz_get_print_state(stz_ctx *ctx, double out[3]);
int stz_set_print_state(stz_ctx *ctx, double strainhist_last);

/* Kernel timing with HIP events on the context stream (measurement hook; no counterpart
   in the reference).  Groups: 0 rates (k_rates), 1 stress solve (all FFT/Green kernels),
   2 selection, 3 transform apply, 4 athermal scan.  ms[g] = summed milliseconds,
   n[g] = launches timed. */
int stz_timing_enable(stz_ctx *ctx, int on);
int stz_timing_get(stz_ctx *ctx, double ms[8], long long n[8]);
/* per-kernel timing inside the stress solve: k_fwd_x, k_ypass(fwd), k_fused, k_ypass(inv), k_inv_x */
int stz_timing_solve_get(stz_ctx *ctx, double ms[8], long long n[8]);

/* Device pointers of the fields stz_rates_noalloc reads (FP64, layouts as stz_upload:
   SIGMA [C][N], Q0 [M][N], DEPS [M][S][N], DEPS_LAST [S][N], PERM/TEMP/TLAST [N]). */
typedef struct stz_rate_fields {
    const double *sigma, *q0, *deps, *deps_last, *perm, *temp, *tlast;
} stz_rate_fields;

/* stz_rates without a context and without allocating: SofteningFactor + ActEnergy + the
   Q<0 count + the KMC rate total over the caller's device fields at clock `totime`, with
   the constants of *p (FP64 glass: p->precision must be STZ_FP64; no crystal-like MRO).
   No counterpart in the reference: it is for callers that cannot call hipMalloc (code
   under HIP graph capture, serving loops that own their buffers).  Kernels are enqueued
   on `stream` (a hipStream_t; NULL = the null stream) and the results are written to
   the caller's device memory (*total_rate, *n_negative) when they run; the call neither
   allocates nor synchronises.  Its block and group rate sums live in one fixed device
   array of 16 MiB per GPU, held until the process ends: modes * ceil(voxels / 256) must
   not exceed 2^21 (e.g. 20 modes up to 6.7 M voxels), else STZ_E_TOO_LARGE.  Calls that
   overlap on one GPU share that array and race; R and the count equal stz_rates' bit for
   bit on the same fields. */
int stz_rates_noalloc(const stz_params *p, const stz_rate_fields *f, double totime, void *stream, double *total_rate,
                      long long *n_negative);

#ifdef __cplusplus
}
#endif
#endif
