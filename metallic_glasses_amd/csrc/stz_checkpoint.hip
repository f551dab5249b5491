// stz_checkpoint.hip -- checkpoint and restart of a run.
//
// The reference saves no restartable state: its outputs are post-processing fields at
// print steps (STZ/bmg.c:238-330), and the MT19937 state, Q0, the transformation strains
// and the softening state are never written, so a run cannot resume.  A checkpoint here
// holds everything the next step reads: the generator (624 words + position), every
// per-voxel array in its storage precision (bit for bit), EpsAvg, <sigma>, the clock,
// the pending uniform stress increment, the print/termination state, the counters and
// the event-local update's bookkeeping, the local temperature and the heat-conduction
// switch, the crystal-like MRO map with the initial barriers a print step restores, and
// per-voxel moduli when they are set (the reference keeps them fixed for a run,
// STZ/init.c:457-465).  Derived arrays (the Q0*exp(-D*S) cache, min Q, block sums) are
// rebuilt by the next rate pass.  A run restarted from a checkpoint continues exactly as
// the saved run does.
#include "stz_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

constexpr char kMagic[8] = {'S', 'T', 'Z', 'C', 'K', 'P', 'T', '1'};

struct CkptHeader {
    char magic[8];
    int32_t version;
    int32_t dim, n0, n1, n2, modes, precision, C, S, rsz, inhom;
    int64_t N;
};

struct CkptScalars {
    double epsavg[6], stress[6], soff[6], ebase[6], usum[6];
    double totime, pend_delta, strainhist_last, soc_elast, endstrain;
    int64_t step, astep, kstep, estep, ntrans, nsolve, avlan, nlocal, rng_count, nmro;
    int32_t pend_comp, printperiod, rng_pos, heat, print_on, pad;
    uint32_t rng_state[624];
};

constexpr size_t kChunk = (size_t)64 << 20;   // staging bytes per transfer

struct Field {
    void *dev;
    size_t bytes;
};

std::vector<Field> fields_of(stz_ctx *c)
{
    const size_t N = (size_t)c->N, R = c->rsz;
    std::vector<Field> f = {
        {c->eps0, (size_t)c->C * N * 8},
        {c->sigma, (size_t)c->C * N * R},
        {c->q0, (size_t)c->M * N * R},
        {c->deps, (size_t)c->M * c->S * N * R},
        {c->deps_last, (size_t)c->S * N * R},
        {c->perm, N * 8},
        {c->temp, N * 8},
        {c->tlast, N * 8},
        {c->tloc, N * 8},
    };
    if (c->q1) f.push_back({c->q1, (size_t)c->M * N * R});   // initial barriers of an MRO context
    if (c->lamx) {
        f.push_back({c->lamx, N * 8});
        f.push_back({c->mux, N * 8});
    }
    return f;
}

CkptHeader header_of(const stz_ctx *c)
{
    CkptHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = 3;
    h.dim = c->dim;
    h.n0 = c->n[0];
    h.n1 = c->n[1];
    h.n2 = c->n[2];
    h.modes = c->M;
    h.precision = c->prec;
    h.C = c->C;
    h.S = c->S;
    h.rsz = (int32_t)c->rsz;
    h.inhom = c->lamx ? 1 : 0;
    h.N = c->N;
    return h;
}

int fail(stz_ctx *c, FILE *f, const char *msg, int rc)
{
    if (f) std::fclose(f);
    c->err = msg;
    return rc;
}

}  // namespace

extern "C" {

int stz_checkpoint_save(stz_ctx *ctx, const char *path)
{
    if (!ctx || !path) return STZ_E_ARG;
    stz_rng_sync(ctx);                      // the host generator at the resident stream position
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, nullptr, "stream synchronisation failed", STZ_E_HIP);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(ctx, nullptr, "cannot open the checkpoint file for writing", STZ_E_ARG);
    const CkptHeader h = header_of(ctx);
    CkptScalars s;
    std::memset(&s, 0, sizeof(s));
    for (int c = 0; c < 6; ++c) {
        s.epsavg[c] = ctx->epsavg[c];
        s.stress[c] = ctx->stress[c];
        s.soff[c] = ctx->soff[c];
        s.ebase[c] = ctx->ebase[c];
        s.usum[c] = ctx->usum[c];
    }
    s.totime = ctx->totime;
    s.pend_delta = ctx->pend_delta;
    s.strainhist_last = ctx->strainhist_last;
    s.soc_elast = ctx->soc_elast;
    s.endstrain = ctx->endstrain;
    s.step = ctx->step;
    s.astep = ctx->astep;
    s.kstep = ctx->kstep;
    s.estep = ctx->estep;
    s.ntrans = ctx->ntrans;
    s.nsolve = ctx->nsolve;
    s.avlan = ctx->avlan;
    s.nlocal = ctx->nlocal;
    s.rng_count = ctx->rng.count;
    s.pend_comp = ctx->pend_comp;
    s.printperiod = ctx->printperiod;
    s.print_on = ctx->print_on;
    s.rng_pos = ctx->rng.pos;
    s.heat = ctx->heat;
    s.nmro = ctx->nmro;
    std::memcpy(s.rng_state, ctx->rng.state, sizeof(s.rng_state));
    if (std::fwrite(&h, sizeof(h), 1, f) != 1 || std::fwrite(&s, sizeof(s), 1, f) != 1)
        return fail(ctx, f, "checkpoint write failed", STZ_E_ARG);
    if (ctx->nmro && std::fwrite(ctx->mro_h.data(), 1, (size_t)ctx->N, f) != (size_t)ctx->N)
        return fail(ctx, f, "checkpoint write failed", STZ_E_ARG);
    std::vector<char> buf(kChunk);
    for (const Field &fd : fields_of(ctx)) {
        for (size_t off = 0; off < fd.bytes; off += kChunk) {
            const size_t n = std::min(kChunk, fd.bytes - off);
            if (hipMemcpy(buf.data(), (const char *)fd.dev + off, n, hipMemcpyDeviceToHost) != hipSuccess)
                return fail(ctx, f, "device-to-host copy failed", STZ_E_HIP);
            if (std::fwrite(buf.data(), 1, n, f) != n) return fail(ctx, f, "checkpoint write failed", STZ_E_ARG);
        }
    }
    if (std::fclose(f) != 0) return fail(ctx, nullptr, "checkpoint close failed", STZ_E_ARG);
    return STZ_OK;
}

int stz_checkpoint_load(stz_ctx *ctx, const char *path)
{
    if (!ctx || !path) return STZ_E_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(ctx, nullptr, "cannot open the checkpoint file", STZ_E_ARG);
    CkptHeader h;
    CkptScalars s;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 3)
        return fail(ctx, f, "not a checkpoint file of this library", STZ_E_ARG);
    const CkptHeader mine = header_of(ctx);
    if (std::memcmp(&h, &mine, sizeof(h)) != 0)
        return fail(ctx, f,
                    "the checkpoint was written by a context of another grid, mode count, precision or "
                    "moduli kind (uniform or per voxel: call stz_set_moduli first)",
                    STZ_E_ARG);
    if (std::fread(&s, sizeof(s), 1, f) != 1 || s.rng_pos < 0 || s.rng_pos > 624 || s.pend_comp < -1 ||
        s.pend_comp >= ctx->C)
        return fail(ctx, f, "truncated or corrupt checkpoint", STZ_E_ARG);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, f, "stream synchronisation failed", STZ_E_HIP);
    // crystal-like MRO map (static after initialisation)
    if (s.nmro < 0 || s.nmro > ctx->N) return fail(ctx, f, "truncated or corrupt checkpoint", STZ_E_ARG);
    ctx->mro_h.assign((size_t)ctx->N, 0);
    if (s.nmro && std::fread(ctx->mro_h.data(), 1, (size_t)ctx->N, f) != (size_t)ctx->N)
        return fail(ctx, f, "truncated checkpoint", STZ_E_ARG);
    {
        const int rc = stz_set_mro_map(ctx);
        if (rc) return fail(ctx, f, "cannot restore the MRO map", rc);
    }
    std::vector<char> buf(kChunk);
    for (const Field &fd : fields_of(ctx)) {
        for (size_t off = 0; off < fd.bytes; off += kChunk) {
            const size_t n = std::min(kChunk, fd.bytes - off);
            if (std::fread(buf.data(), 1, n, f) != n) return fail(ctx, f, "truncated checkpoint", STZ_E_ARG);
            if (hipMemcpy((char *)fd.dev + off, buf.data(), n, hipMemcpyHostToDevice) != hipSuccess)
                return fail(ctx, f, "host-to-device copy failed", STZ_E_HIP);
        }
    }
    std::fclose(f);
    for (int c = 0; c < 6; ++c) {
        ctx->epsavg[c] = s.epsavg[c];
        ctx->stress[c] = s.stress[c];
        ctx->soff[c] = s.soff[c];
        ctx->ebase[c] = s.ebase[c];
        ctx->usum[c] = s.usum[c];
    }
    ctx->totime = s.totime;
    ctx->pend_delta = s.pend_delta;
    ctx->pend_comp = s.pend_comp;
    ctx->strainhist_last = s.strainhist_last;
    ctx->soc_elast = s.soc_elast;
    ctx->endstrain = s.endstrain;
    ctx->printperiod = s.printperiod;
    ctx->print_on = s.print_on != 0;
    ctx->step = s.step;
    ctx->astep = s.astep;
    ctx->kstep = s.kstep;
    ctx->estep = s.estep;
    ctx->ntrans = s.ntrans;
    ctx->nsolve = s.nsolve;
    ctx->avlan = s.avlan;
    ctx->nlocal = s.nlocal;
    std::memcpy(ctx->rng.state, s.rng_state, sizeof(s.rng_state));
    ctx->rng.pos = s.rng_pos;
    ctx->rng.count = s.rng_count;
    ctx->heat = s.heat != 0;
    stz_rng_replaced(ctx);                  // resident lookup windows are rebuilt from here
    return stz_invalidate_derived(ctx);
}

}  // extern "C"
