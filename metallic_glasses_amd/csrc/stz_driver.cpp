// stz_driver.cpp -- the reference's main program (STZ/bmg.c) at one rank, on the C-ABI.
//
// Reads the reference's input file (STZ/io.c:4-101, labels matched literally), plus an
// optional extension block after the last reference field, runs the main loop
// (STZ/bmg.c:138-411) device-resident through stz_run_steps (stopping at print steps)
// and stz_step_print, and writes the reference's
// outputs: actQ.out (per thermal event), SS.out, SOC.out, ElastSOC.out,
// ElasticEnergy.out, EA.out, EpsAvg.out, avg.out, and the Strain%03d_P00.out /
// Stress%03d_P00.out field dumps and the X1R/X2R (X3R in 3-D) displacement records at
// print steps.
//
// usage: stz_run [input-file]      (default "input", as STZ/bmg.c:31)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "stz.h"

namespace {

struct Input {
    double E = 0, nu = 0, q0amp = 0, strainrate = 0, strainincrement = 0, randomratio = 0;
    int seed = 0, printperiod = 0, softening = 0, initmro = 0;
    double endstrain = 0, relxtime = 0, xtalshear = 0, xtalq0 = 0, resiratio = 0, mrofrac = 0, xq0 = 0;
    char resfile[256] = {0};
    // extension block
    int dim = 2, n[3] = {512, 512, 1}, loading = STZ_TENSION, precision = STZ_FP64, modes = 20;
    int athermal_correction = -1, field_output = 1;
    double gamma0 = 0.14, temperature = 300.0, q0_std = -1.0;
};

// label-matched reads with the reference's messages (STZ/io.c:7-97)
template <typename... A>
void want(FILE *f, int n, const char *fmt, const char *what, A... a)
{
    if (std::fscanf(f, fmt, a...) != n) std::fprintf(stderr, "ReadInput: Could not match %s\n", what);
}

void read_input(FILE *f, Input &in)
{
    want(f, 2, "Moduli(E, nu):\n%lf %lf\n\n", "Moduli", &in.E, &in.nu);
    want(f, 1, "Stress-free energy barrier:\n%lf\n\n", "Stress-free energy barrier", &in.q0amp);
    want(f, 1, "Applied strain rate:\n%lf\n\n", "Applied strain rate", &in.strainrate);
    want(f, 1, "Strain increment:\n%lf\n\n", "Strain increment", &in.strainincrement);
    want(f, 1, "Random seed:\n%d\n\n", "Random seed", &in.seed);
    want(f, 1, "Random ratio of fluctuation in energy barrier:\n%lf\n\n",
         "Random ratio of fluctuation in energy barrier", &in.randomratio);
    want(f, 1, "Print period:\n%d\n\n", "Print period", &in.printperiod);
    want(f, 1, "End Strain:\n%lf\n\n", "EndStrain", &in.endstrain);
    want(f, 1, "Softening Flag:\n%d\n\n", "Softening Flag", &in.softening);
    want(f, 1, "Relaxation Time:\n%lf\n\n", "Relaxation Time", &in.relxtime);
    want(f, 1, "Characteristic Xtal shear:\n%lf\n\n", "Characteristic xtal shear", &in.xtalshear);
    want(f, 1, "Barrier for Xtal shear:\n%lf\n\n", "Barrier for xtal shear", &in.xtalq0);
    want(f, 1, "Ratio for generating residual stress:\n%lf\n\n", "Residual stress Ratio", &in.resiratio);
    want(f, 1, "Residual stress data file:\n%255s\n\n", "Residual stress data file", in.resfile);
    want(f, 1, "Volume fraction of crystal-like MRO:\n%lf\n\n", "volume fraction of crystal-like MRO", &in.mrofrac);
    want(f, 1, "Initialization condition for MRO:\n%d\n\n", "initialization condition for MRO", &in.initmro);
    want(f, 1, "Stress-free energy barrier for crystal-like MRO:\n%lf\n\n",
         "Stress-free energy barrier for crystal-like MRO", &in.xq0);
    // optional extension block (absent in reference inputs): "Label:\nvalue(s)\n\n"
    char label[128];
    while (std::fscanf(f, " %127[^:\n]:", label) == 1) {
        std::string L(label);
        if (L == "Dimension") std::fscanf(f, "%d", &in.dim);
        else if (L == "Grid") std::fscanf(f, "%d %d %d", &in.n[0], &in.n[1], &in.n[2]);
        else if (L == "Loading") {
            char w[32];
            if (std::fscanf(f, "%31s", w) == 1) in.loading = (std::strcmp(w, "shear") == 0) ? STZ_SHEAR : STZ_TENSION;
        } else if (L == "Precision") {
            char w[32];
            if (std::fscanf(f, "%31s", w) == 1) in.precision = (std::strcmp(w, "mixed") == 0) ? STZ_MIXED : STZ_FP64;
        } else if (L == "Modes") std::fscanf(f, "%d", &in.modes);
        else if (L == "Transformation strain") std::fscanf(f, "%lf", &in.gamma0);
        else if (L == "Temperature") std::fscanf(f, "%lf", &in.temperature);
        else if (L == "Barrier std") std::fscanf(f, "%lf", &in.q0_std);
        else if (L == "Athermal correction") std::fscanf(f, "%d", &in.athermal_correction);
        else if (L == "Field output") std::fscanf(f, "%d", &in.field_output);
        else {
            std::fprintf(stderr, "ReadInput: unknown extension field '%s'\n", label);
            char rest[256];
            if (!std::fgets(rest, sizeof rest, f)) break;
        }
    }
}

struct Soc {
    // countSOC[0] avalanche size, countSOC[1] number of avalanches (STZ/bmg.c:23-28)
    std::vector<int> size{1}, count{0};
    std::vector<double> etime, eval;
    long long counter = 0;
    double elast = 0.0;
    // UpdateSOC, STZ/evolution.c:5-45
    void update(double totime)
    {
        const long long a = counter;
        if (a > (long long)size.size()) {
            for (long long i = (long long)size.size(); i < a; ++i) {
                size.push_back((int)(i + 1));
                count.push_back(0);
            }
            count[a - 1] = 1;
        } else if (a > 0) {
            count[a - 1] += 1;
        }
        counter = 0;
        if (std::fabs(elast) > 1e-17) {
            etime.push_back(totime);
            eval.push_back(elast);
            elast = 0.0;
        }
    }
};

void write_field(const char *name, int idx, const std::vector<double> &v, int n0, long long rows)
{
    char fname[64];
    std::snprintf(fname, sizeof fname, "%s%03d_P%02d.out", name, idx, 0);
    FILE *f = std::fopen(fname, "w");
    if (!f) {
        std::perror("Error opening file\n");
        std::exit(1);
    }
    for (long long r = 0; r < rows; ++r) {
        for (int i = 0; i < n0; ++i) std::fprintf(f, "%lf ", v[(size_t)(r * n0 + i)]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void write_summaries(const std::vector<double> &sh, const std::vector<double> &st, const std::vector<double> &eh,
                     const Soc &soc, const double *epsavg, int C)
{
    FILE *f = std::fopen("SS.out", "w");
    for (size_t i = 0; i < sh.size(); ++i) std::fprintf(f, "%.6e\t%.6e\n", sh[i], st[i]);
    std::fclose(f);
    f = std::fopen("SOC.out", "w");
    for (size_t i = 0; i < soc.size.size(); ++i) std::fprintf(f, "%d\t%d\n", soc.size[i], soc.count[i]);
    std::fclose(f);
    f = std::fopen("ElastSOC.out", "w");
    for (size_t i = 0; i < soc.etime.size(); ++i) std::fprintf(f, "%.6e\t%.6e\n", soc.etime[i], soc.eval[i]);
    std::fclose(f);
    f = std::fopen("ElasticEnergy.out", "w");
    for (size_t i = 0; i < sh.size(); ++i) std::fprintf(f, "%.6e\t%.6e\n", sh[i], eh[i]);
    std::fclose(f);
    f = std::fopen("EA.out", "w");
    for (int c = 0; c < C; ++c) std::fprintf(f, c + 1 < C ? "%.6e\t" : "%.6e\n", epsavg[c]);
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv)
{
    const char *path = argc > 1 ? argv[1] : "input";
    FILE *fin = std::fopen(path, "r");
    if (!fin) {
        std::perror("input");
        return 1;
    }
    Input in;
    read_input(fin, in);
    std::fclose(fin);
    std::printf("Input file read\n");

    stz_params p;
    std::memset(&p, 0, sizeof p);
    p.dim = in.dim;
    p.n[0] = in.n[0];
    p.n[1] = in.n[1];
    p.n[2] = in.dim == 2 ? 1 : in.n[2];
    p.modes = in.modes;
    p.precision = in.precision;
    p.loading = in.loading;
    p.athermal_correction = in.athermal_correction >= 0 ? in.athermal_correction : (in.dim == 2 ? 0 : 1);
    p.softening = in.softening;
    p.seed = in.seed;
    p.E = in.E;
    p.nu = in.nu;
    p.double_nu = in.dim == 2 ? 1 : 0;                 // LOW_T, STZ/init.c:26
    p.q0_mean = in.q0amp;
    p.q0_std = in.q0_std > 0 ? in.q0_std : 0.25;      // STZ/init.c:185
    p.gamma0 = in.gamma0;
    p.temperature = in.temperature;
    p.trial_frequency = 1e13;
    p.lx = 0.7;
    p.ly = 0.7;
    p.lz = 1.0;
    p.strainrate = in.strainrate;
    p.strainincrement = in.strainincrement;
    p.perm_soft = 13;
    p.temp_soft = 39;
    p.perm_cap = -std::log(0.6);
    p.temp_soft_barrier = 0.37;
    p.athermal_threshold = 0.33;
    p.mro_fraction = in.mrofrac;
    p.mro_q0 = in.xq0;
    p.random_ratio = in.randomratio;
    if (in.initmro != 0 && in.initmro != 1) {
        std::fprintf(stderr, "Error: wrong option for MRO initialization\n");   // STZ/init.c:433-435
        return 1;
    }

    stz_ctx *ctx = nullptr;
    int rc = stz_create(&p, &ctx);
    if (rc) {
        std::fprintf(stderr, "stz_create failed (%d)\n", rc);
        return 1;
    }
    if (in.initmro == 1) {
        // MROstructure.ms: one line "i j type theta" per voxel, 1-based, slow index first,
        // in the voxel order (STZ/init.c:398-419)
        const long long NV = (long long)p.n[0] * p.n[1] * p.n[2];
        std::vector<int> type((size_t)NV);
        std::vector<double> theta((size_t)NV);
        FILE *fm = std::fopen("MROstructure.ms", "r");
        if (!fm) {
            std::perror("MROstructure.ms");
            return 1;
        }
        char line[256];
        for (long long m = 0; m < NV; ++m) {
            int ii = 0, jj = 0;
            if (!std::fgets(line, sizeof line, fm) ||
                std::sscanf(line, "%d %d %d %lf", &ii, &jj, &type[(size_t)m], &theta[(size_t)m]) != 4 ||
                (long long)(ii - 1) * p.n[0] + (jj - 1) != m) {
                std::fprintf(stderr, "Error in reading MROstructure.ms file (inconsistent index system)!!\n");
                return 1;
            }
        }
        std::fclose(fm);
        rc = stz_init_state_mro(ctx, type.data(), theta.data());
    } else {
        rc = stz_init_state(ctx);
    }
    if (rc) {
        std::fprintf(stderr, "stz_init_state failed (%d): %s\n", rc, stz_last_error(ctx));
        return 1;
    }
    stz_set_print(ctx, in.printperiod, in.endstrain);
    const int C = in.dim == 2 ? 3 : 6;
    const int lc = in.loading == STZ_TENSION ? 0 : (C == 3 ? 2 : 5);
    const long long N = (long long)p.n[0] * p.n[1] * p.n[2];
    const long long rows = (long long)p.n[1] * p.n[2];

    double epsavg[6], stress[6], totime;
    stz_get_scalars(ctx, epsavg, stress, &totime);
    std::printf("PE#0:step=0\ttime=%lf\tstrain=[%.3e, %.3e, %.3e]\nstress=[%.3e, %.3e, %.3e]\n\n", totime, epsavg[0],
                epsavg[1], epsavg[C - 1], stress[0], stress[1], stress[C - 1]);

    // the MRO map: Initial_MROstructure_S000.iout, native int32 in voxel order
    // (WriteMROMPI_INT, STZ/io.c:321-347, STZ/init.c:437), and the actQ.out type column
    std::vector<double> mrod((size_t)N);
    std::vector<int> mro((size_t)N);
    stz_download(ctx, STZ_F_MRO, mrod.data());
    for (long long v = 0; v < N; ++v) mro[(size_t)v] = (int)mrod[(size_t)v];
    {
        FILE *fo = std::fopen("Initial_MROstructure_S000.iout", "wb");
        if (fo) {
            std::fwrite(mro.data(), sizeof(int), (size_t)N, fo);
            std::fclose(fo);
        }
    }
    FILE *fq = std::fopen("actQ.out", "w");
    std::vector<double> strainhist{0.0}, stresshist{0.0}, elhist{0.0};
    Soc soc;
    std::vector<double> eps0((size_t)(C * N)), sig((size_t)(C * N)), eq((size_t)N), disp((size_t)(in.dim * N));
    long long step = 0, astep = 0, kstep = 0, estep = 0;
    double prev_time = totime;
    // the main loop runs device-resident in batches; each batch returns after a print
    // step, before its re-solve, so the print block sees the reference's state
    stz_run_stop_at_print(ctx, 1);
    std::vector<stz_step_result> recs(4096);
    stz_step_result r;
    std::memset(&r, 0, sizeof(r));
    bool finished = false;
    while (!finished) {
        int done = 0;
        const int brc = stz_run_steps(ctx, (int)recs.size(), recs.data(), &done);
        if (brc < 0) {
            std::fprintf(stderr, "step %lld failed (%d): %s\n", step + done + 1, brc, stz_last_error(ctx));
            return 2;
        }
        for (int i = 0; i < done && !finished; ++i) {
            r = recs[(size_t)i];
            rc = (i + 1 == done && !r.printed) ? brc : 0;
            ++step;
            if (r.kind == STZ_STEP_ATHERMAL) {
                ++astep;
                soc.counter += r.n_transformed;
                soc.elast += r.soc_elast;
            } else {
                soc.update(prev_time);                      // UpdateSOC before KMC, STZ/bmg.c:167
                if (r.kind == STZ_STEP_THERMAL) {
                    ++kstep;
                    std::fprintf(fq, "%d\t%.6e\n", mro[(size_t)r.site], r.q_chosen);   // MRO type and Q, STZ/evolution.c:243
                } else {
                    ++estep;
                }
            }
            prev_time = r.totime;
            if (r.printed) {
                std::fflush(fq);
                double el = 0.0;
                stz_elastic_energy(ctx, &el);               // STZ/bmg.c:174-178, before the re-solve
                strainhist.push_back(r.epsavg[lc]);
                stresshist.push_back(r.stress[lc]);
                elhist.push_back(el);
                const int np = (int)strainhist.size();
                std::printf("PE#0:step=%lld\ttime=%lf\tstrain=[%.3e, %.3e, %.3e]\nstress=[%.3e, %.3e, %.3e]\n", step,
                            r.totime, r.epsavg[0], r.epsavg[1], r.epsavg[lc], r.stress[0], r.stress[1], r.stress[lc]);
                std::printf("step=%lld\nastep=%lld\nkstep=%lld\nestep=%lld\nN_print=%d\n\n", step, astep, kstep, estep, np);
                stz_download(ctx, STZ_F_EPS0, eps0.data());
                double avg[6] = {0, 0, 0, 0, 0, 0};
                for (int c = 0; c < C; ++c) {
                    for (long long v = 0; v < N; ++v) avg[c] += eps0[(size_t)(c * N + v)];
                    avg[c] /= (double)N;
                }
                if (in.field_output) {
                    // von Mises-like magnitudes of STZ/io.c:260-273: sqrt(xx^2 + xy^2)
                    stz_download(ctx, STZ_F_SIGMA, sig.data());
                    const int xy = C == 3 ? 2 : 5;
                    for (long long v = 0; v < N; ++v)
                        eq[(size_t)v] = std::sqrt(eps0[(size_t)v] * eps0[(size_t)v] +
                                                  eps0[(size_t)(xy * N + v)] * eps0[(size_t)(xy * N + v)]);
                    write_field("Strain", np - 1, eq, p.n[0], rows);
                    for (long long v = 0; v < N; ++v)
                        eq[(size_t)v] = std::sqrt(sig[(size_t)v] * sig[(size_t)v] +
                                                  sig[(size_t)(xy * N + v)] * sig[(size_t)(xy * N + v)]);
                    write_field("Stress", np - 1, eq, p.n[0], rows);
                }
                rc = stz_step_print(ctx, &r);               // re-solve, STZ/bmg.c:289-295
                if (rc < 0) {
                    std::fprintf(stderr, "print re-solve failed (%d): %s\n", rc, stz_last_error(ctx));
                    return 2;
                }
                if (in.field_output) {
                    // displacement record X1R/X2R (X3R in 3-D), STZ/bmg.c:286-330
                    const int drc = stz_displacement(ctx, disp.data());
                    if (drc < 0) {
                        std::fprintf(stderr, "displacement failed (%d): %s\n", drc, stz_last_error(ctx));
                        return 2;
                    }
                    static const char *names[3] = {"X1R", "X2R", "X3R"};
                    for (int a = 0; a < in.dim; ++a) {
                        std::vector<double> comp(disp.begin() + (size_t)a * N, disp.begin() + (size_t)(a + 1) * N);
                        write_field(names[a], np - 1, comp, p.n[0], rows);
                    }
                }
                write_summaries(strainhist, stresshist, elhist, soc, r.epsavg, C);
                std::printf("step=%lld\nastep=%lld\nkstep=%lld\nestep=%lld\n", step, astep, kstep, estep);
                FILE *f = std::fopen("EpsAvg.out", "a");
                for (int c = 0; c < C; ++c) std::fprintf(f, c + 1 < C ? "%lf " : "%lf\n", r.epsavg[c]);
                std::fclose(f);
                f = std::fopen("avg.out", "a");
                for (int c = 0; c < C; ++c) std::fprintf(f, c + 1 < C ? "%lf " : "%lf\n", avg[c]);
                std::fclose(f);
            }
            if (rc == 1) {
                if (std::fabs(r.epsavg[lc]) > in.endstrain) std::printf("PE#0: macro strain exceeds %lf\n", in.endstrain);
                else std::printf("\n\nUnstable system! Solution is divergent (occilating)\n");
                finished = true;
            }
        }
        std::fflush(fq);
    }
    write_summaries(strainhist, stresshist, elhist, soc, r.epsavg, C);
    std::printf("step=%lld\nastep=%lld\nkstep=%lld\nestep=%lld\n", step, astep, kstep, estep);
    std::fclose(fq);
    stz_destroy(ctx);
    std::printf("\nProgram finished\n");
    return 0;
}
