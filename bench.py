#!/usr/bin/env python3
"""kMC events/s of the STZ kinetic Monte Carlo step on one MI355X.

Workload (default): 128x128x128 voxels, FP64 plain GPU path, mu = 35 GPa, nu = 0.35,
20 shear-transformation modes, gamma0 = 0.1, Q0 ~ N(2.0, 0.2) eV from seed 1234,
T = 300 K, simple shear at 1e-4/s.  A step is one iteration of the reference main loop
(rates over all modes and voxels, then a thermal event with selection, eigenstrain
update and a full Green's-function stress solve, or an athermal avalanche, or a pure
elastic increment).  Warm-up runs the elastic loading until W thermal events have
happened; the timed region is exactly K steps; `value` = thermal events / wall time.

Also reported: the mixed-precision path (FP32 FFT/Green/rates, FP64 sums, clock,
eigenstrain and mean stress) from the same state with its stress-field error against
FP64; the roofline of the dominant kernel (k_rates, HBM-bound) from HIP events on the
library's stream; the CPU FP64 oracle timed on one host core on a bounded sample that
starts from the GPU's timed-region state.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--grid", default="128,128,128")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-replicas", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--ensemble-warmup", type=int, default=200)
    ap.add_argument("--pmc-target", default=None,
                    help="comma list of rates,copy: run only those kernels (a few launches) for a "
                         "rocprofv3 --pmc pass, then exit")
    ap.add_argument("--pmc-summary", default=None,
                    help="directory of rocprofv3 --pmc output: print per-kernel mean counters, then exit")
    ap.add_argument("--trace-summary", default=None,
                    help="directory of rocprofv3 --kernel-trace output: print per-kernel durations, then exit")
    return ap.parse_args()


COPY_PROBE_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kb_rates_128cube_fp64_traffic.txt")


def pmc_target(a, mg):
    """A short run of the kernels whose byte counters are read: `rates` = the bench
    workload's resident step (every step launches kb_rates once), `copy` = the plain copy
    of COPY_PROBE_BYTES that calibrates the counters."""
    for t in a.pmc_target.split(","):
        if t == "copy":
            ms, gbs = mg.copy_probe(COPY_PROBE_BYTES, 5)
            print(f"# copy probe {COPY_PROBE_BYTES} B: {ms:.4f} ms {gbs:.1f} GB/s", file=sys.stderr, flush=True)
        elif t == "rates":
            dims = tuple(int(x) for x in a.grid.split(","))
            g = mg.Stz(**mg.config_3d(n=dims, seed=a.seed, q0_std=0.2))
            g.set_print(5, 1.0)
            rc, res = g.run_steps(a.steps)
            print(f"# rates target: {len(res)} resident steps of {dims}", file=sys.stderr, flush=True)
            g.close()
        else:
            raise SystemExit(f"unknown pmc target {t}")


def _active(vals):
    """Dispatches that did the kernel's work: a replica that halted (or a step of another
    kind) still launches every kernel of the step, which then exits at once."""
    vals = np.asarray(vals, dtype=np.float64)
    return vals[vals > 0.05 * vals.max()] if len(vals) and vals.max() > 0 else vals


def pmc_summary(d):
    """Per kernel and counter in rocprofv3's counter CSV files: mean over all dispatches
    and over the active ones (value above 5% of the largest)."""
    import csv
    import glob
    acc = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                key = (row["Kernel_Name"], row["Counter_Name"])
                disp = row.get("Dispatch_Id", row.get("Correlation_Id"))
                acc.setdefault(key, {})
                acc[key][disp] = acc[key].get(disp, 0.0) + float(row["Counter_Value"])
    for (k, c), v in sorted(acc.items()):
        vals = np.array(list(v.values()))
        act = _active(vals)
        print(f"{c:12s} dispatches {len(vals):5d} mean {vals.mean():14.1f} active {len(act):5d} "
              f"mean {act.mean() if len(act) else 0.0:14.1f} max {vals.max():14.1f}  {k[:100]}")


def trace_summary(d):
    """Per kernel in rocprofv3's kernel-trace CSV files: calls, mean duration over all
    dispatches and over the active ones (duration above 5% of the longest)."""
    import csv
    import glob
    acc = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                acc.setdefault(row["Kernel_Name"], []).append(
                    (float(row["End_Timestamp"]) - float(row["Start_Timestamp"])) * 1e-3)
    tot = sum(sum(v) for v in acc.values())
    print(f"{'kernel':100s} {'calls':>6s} {'mean_us':>10s} {'active':>6s} {'act_mean_us':>11s} {'pct':>6s}")
    for k, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        vals = np.array(v)
        act = _active(vals)
        print(f"{k[:100]:100s} {len(vals):6d} {vals.mean():10.3f} {len(act):6d} {act.mean():11.3f} "
              f"{100 * vals.sum() / tot:6.2f}")


def traffic_from_profile(dims, precision):
    """HBM bytes per kb_rates launch from the committed rocprofv3 --pmc summary (FETCH_SIZE
    and WRITE_SIZE, each scaled by the copy probe's calibration factor), when that summary
    was measured on this workload; else None."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    kv = {}
    for ln in lines:
        if ln.startswith("traffic_") and ":" in ln:
            k, v = ln.split(":", 1)
            kv[k.strip()] = v.strip()
    if kv.get("traffic_workload") != f"{dims[0]}x{dims[1]}x{dims[2]} {precision}":
        return None
    return float(kv["traffic_bytes_per_launch"])


def rate_bytes_per_voxel(C, M, S, b):
    # sigma (C), Q0 (M), transformation strains (S*M), previous strain (S) in the storage
    # precision, plus Perm, Temp, tlast in FP64
    return (C + M + S * M + S) * b + 3 * 8


def fft_table(dims, C, N, b, sms, sn):
    """Algorithmic bytes and achieved GB/s of the five stress-solve kernels (storage
    precision b bytes; eps0 is always FP64)."""
    h0 = dims[0] // 2 + 1
    spec = C * dims[1] * dims[2] * h0 * 2 * b
    fft_bytes = [C * N * 8 + spec, 2 * spec, 2 * spec, 2 * spec, spec + C * N * 8 + C * N * b]
    names = ["k_fwd_x", "k_ypass_fwd", "k_fused", "k_ypass_inv", "k_inv_x"]
    out = []
    for i in range(5):
        t = sms[i] / max(sn[i], 1)
        gbs = fft_bytes[i] / (t * 1e-3) / 1e9 if t > 0 else 0.0
        out.append({"kernel": names[i], "avg_ms": round(t, 5), "GB/s": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4)})
    return out


def snapshot(g, mg):
    st = {f: g.download(f) for f in (mg.F_EPS0, mg.F_SIGMA, mg.F_Q0, mg.F_DEPS, mg.F_DEPS_LAST,
                                     mg.F_PERM, mg.F_TEMP, mg.F_TLAST)}
    e, s, t = g.scalars()
    return st, (e, s, t), g.rng_state(), g.print_state()


def restore(g, snap, mg):
    st, (e, s, t), rng, pr = snap
    for f, a in st.items():
        g.upload(f, a)
    g.set_scalars(e, s, t)
    g.set_rng_state(rng)
    g.set_print(int(pr[1]), pr[2])
    g.set_print_state(pr[0])


def run_steps(g, k, curve=None):
    t0 = time.perf_counter()
    rc, res = g.run_steps(k)          # device-resident steps; the host synchronises once per batch
    dt = time.perf_counter() - t0
    kinds = np.zeros(3, dtype=np.int64)
    events = []
    for r in res:
        kinds[r.kind] += 1
        if r.kind == 1:
            events.append((r.site, r.mode))
        if curve is not None:
            curve.append((r.kind, r.site, r.mode, r.n_transformed, r.epsavg[5], r.stress[5]))
    return dt, kinds, events


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.pmc_summary:
        pmc_summary(a.pmc_summary)
        return
    if a.trace_summary:
        trace_summary(a.trace_summary)
        return
    import metallic_glasses_amd as mg
    if a.pmc_target:
        pmc_target(a, mg)
        return

    dims = tuple(int(x) for x in a.grid.split(","))
    kw = mg.config_3d(n=dims, seed=a.seed, q0_std=0.2)
    t_init = time.perf_counter()
    g = mg.Stz(**kw)
    g.set_print(5, 1.0)
    init_s = time.perf_counter() - t_init
    tune64 = g.autotune(reps=10)
    print(f"# autotuned solve tiles (x fwd rows, y cols, fused cols, x inv rows): {tune64[0]} "
          f"ms {np.round(tune64[1], 4).tolist()}", file=sys.stderr, flush=True)
    print(f"# init {dims} fp64: {init_s:.1f}s", file=sys.stderr, flush=True)

    # warm-up: elastic loading until the first thermal events, then W events
    nev = 0
    nst = 0
    while nev < a.warmup:
        rc, r = g.step()
        nst += 1
        nev += r.kind == 1
        if rc:
            break
    g.run_steps(20)      # resident path warm (lookup window built) before the timed region
    print(f"# warm-up: {nst} + 20 steps, {nev} thermal events", file=sys.stderr, flush=True)

    snap = None
    if not a.no_cpu or not a.no_mixed:
        snap = snapshot(g, mg)

    # timed region: device-resident steps replayed from a hipGraph
    curve64 = []
    dt, kinds, events = run_steps(g, a.steps, curve64)
    sig64_end = g.download(mg.F_SIGMA) if not a.no_mixed else None
    N = g.N
    C, M, S = g.C, g.M, g.S
    ev_fp64 = kinds[1] / dt
    _, copy_gbs = mg.copy_probe(COPY_PROBE_BYTES, 10)
    # kernel timing with HIP events on the library's stream, continuing the same run:
    # the resident step in direct launches (events around every rate pass), then the
    # per-call step (events around every solve kernel)
    g.run_config(use_graph=0)
    g.timing(True)
    t0 = time.perf_counter()
    _, tk, _ = run_steps(g, a.timing_steps)
    dt_direct = time.perf_counter() - t0
    ms, n, _, _ = g.timing_get()
    k4_ms = ms[0] / max(n[0], 1)
    g.timing(False)
    g.timing(True)
    for _ in range(a.timing_steps // 4):
        rc_, _r = g.step()
        if rc_:
            break
    ms2, n2, sms, sn = g.timing_get()
    g.timing(False)
    k4_bytes = rate_bytes_per_voxel(C, M, S, 8) * N
    achieved = k4_bytes / (k4_ms * 1e-3) / 1e9
    solve_ms = ms2[1] / max(n2[1], 1)
    print(f"# fp64: {a.steps} steps in {dt:.3f}s kinds(el,th,ath)={kinds.tolist()} "
          f"k_rates {k4_ms:.3f} ms ({achieved:.0f} GB/s) solve {solve_ms:.3f} ms "
          f"per-kernel {np.round(sms[:5] / np.maximum(sn[:5], 1), 4).tolist()}; direct launches "
          f"{tk[1] / dt_direct:.1f} events/s", file=sys.stderr, flush=True)

    out = {
        "metric": "kMC events/s on 1 GPU, mixed vs FP64 at fixed stress err; FFT HBM roofline %",
        "value": round(float(ev_fp64), 3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3 / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded MT19937 initial state, as the reference)",
        "config": {"workload": f"STZ kMC {dims[0]}x{dims[1]}x{dims[2]} FP64, M=20, simple shear, seed {a.seed}",
                   "grid": list(dims), "modes": M, "precision": "fp64", "parallelism": "single GPU"},
        "steps_by_kind": {"elastic": int(kinds[0]), "thermal": int(kinds[1]), "athermal": int(kinds[2])},
        "events_per_s_direct_launches": round(float(tk[1] / dt_direct), 3),
        "roofline": {"bound": "hbm", "kernel": "kb_rates", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(dims, "fp64"),
                     "bytes_per_launch": int(k4_bytes), "avg_ms": round(k4_ms, 5)},
        "stream_copy_GBs": round(copy_gbs, 1),
        "solve_ms": round(solve_ms, 5),
        "solve_tiles": tune64[0],
    }
    fft = fft_table(dims, C, N, 8, sms, sn)
    out["fft_roofline"] = fft

    if not a.no_mixed and snap is not None:
        kw_m = dict(kw, precision=mg.MIXED)
        ref = mg.Stz(init_state=False, **kw)
        restore(ref, snap, mg)
        ref.stress_solve()
        sref = ref.download(mg.F_SIGMA)
        ref.close()
        gm = mg.Stz(init_state=False, **kw_m)
        restore(gm, snap, mg)
        tune_m = gm.autotune(reps=10)
        gm.stress_solve()
        smix = gm.download(mg.F_SIGMA)
        err = float(np.linalg.norm(smix - sref) / np.linalg.norm(sref))
        restore(gm, snap, mg)
        curvem = []
        dtm, kinds_m, events_m = run_steps(gm, a.steps, curvem)
        smix_end = gm.download(mg.F_SIGMA)
        gm.run_config(use_graph=0)
        gm.timing(True)
        run_steps(gm, a.timing_steps)
        msm, nm, _, _ = gm.timing_get()
        gm.timing(False)
        k4m = msm[0] / max(nm[0], 1)
        gm.timing(True)
        for _ in range(a.timing_steps // 4):
            rc_, _r = gm.step()
            if rc_:
                break
        _, _, smsm, snm = gm.timing_get()
        gm.timing(False)
        kb = rate_bytes_per_voxel(C, M, S, 4) * N
        same = 0
        for x, y in zip(events, events_m):
            if x != y:
                break
            same += 1
        # stress-strain curve (loaded component of <sigma> per step) over the steps whose
        # step sequence is identical in both runs
        nsame = 0
        for x, y in zip(curve64, curvem):
            if x[:4] != y[:4]:
                break
            nsame += 1
        c64 = np.array([x[5] for x in curve64[:nsame]])
        cm = np.array([y[5] for y in curvem[:nsame]])
        ss_err = float(np.max(np.abs(cm - c64)) / max(np.max(np.abs(c64)), 1e-30)) if nsame else None
        end_err = (float(np.linalg.norm(smix_end - sig64_end) / np.linalg.norm(sig64_end))
                   if nsame == len(curve64) == len(curvem) else None)
        out["mixed"] = {"value": round(float(kinds_m[1] / dtm), 3), "unit": "events/s",
                        "stress_rel_err": err, "stress_rel_err_end": end_err, "ss_curve_rel_err": ss_err,
                        "steps_identical": nsame, "ms_per_step": round(dtm * 1e3 / a.steps, 4),
                        "speedup_vs_fp64": round(float(kinds_m[1] / dtm) / ev_fp64, 3),
                        "k_rates_GBs": round(kb / (k4m * 1e-3) / 1e9, 1),
                        "events_identical_to_fp64": same, "solve_tiles": tune_m[0],
                        "fft_roofline": fft_table(dims, C, N, 4, smsm, snm)}
        gm.close()
        print(f"# mixed: {out['mixed']}", file=sys.stderr, flush=True)

    if not a.no_cpu and snap is not None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        cores = min(os.cpu_count() or 1, int(os.environ.get("STZ_CPU_THREADS", "16")))
        cores = po.set_threads(cores)
        o = po.Oracle(empty=True, **kw)
        st, (e, s, t), rng, pr = snap
        for f_g, f_o in [(mg.F_EPS0, po.F_EPS0), (mg.F_SIGMA, po.F_SIGMA), (mg.F_Q0, po.F_Q0),
                         (mg.F_DEPS, po.F_DEPS), (mg.F_DEPS_LAST, po.F_DEPS_LAST), (mg.F_PERM, po.F_PERM),
                         (mg.F_TEMP, po.F_TEMP), (mg.F_TLAST, po.F_TLAST)]:
            o.field(f_o)[:] = st[f_g]
        o.set_scalars(e, s, t)
        r = o.rng()
        for i in range(624):
            r.contents.mt[i] = int(rng[i])
        r.contents.mti = int(rng[624])
        o.set_print(int(pr[1]), pr[2])
        po.lib().ora_set_strainhist(o._h, float(pr[0]))
        cev = 0
        csteps = 0
        cmatch = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < a.cpu_seconds or cev == 0:
            rc, info = o.step()
            csteps += 1
            if info.kind == 1:
                if cev < len(events) and (info.site, info.mode) == events[cev] and cmatch == cev:
                    cmatch += 1
                cev += 1
            if rc:
                break
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cev / cdt, 5), "unit": "events/s", "cores": cores, "kind": "port",
                               "sample": f"{csteps} steps ({cev} thermal events) of the same {dims[0]}^3 FP64 run "
                                         f"from the GPU's timed-region start state, {cdt:.1f}s on {cores} threads",
                               "events_matching_gpu": cmatch}
        print(f"# cpu: {out['cpu_baseline']}", file=sys.stderr, flush=True)

    if not a.no_ensemble:
        out["ensemble"] = ensemble_leg(a, mg)
        print(f"# ensemble: {out['ensemble']}", file=sys.stderr, flush=True)

    print(json.dumps(out))


ENSEMBLE_GRIDS = [(32, 32, 32), (64, 64, 128), (96, 96, 96), (128, 128, 128)]


def ensemble_plan(nrep):
    """Replica r: grid ENSEMBLE_GRIDS[r % 4], seed r, (T_r, std_r) drawn in pairs from
    MT19937 seeded 4242 (numpy's legacy generator): T ~ U[250, 450] K, std ~ U[0.1, 0.3] eV."""
    rs = np.random.RandomState(4242)
    plan = []
    for r in range(nrep):
        T = float(rs.uniform(250.0, 450.0))
        sd = float(rs.uniform(0.1, 0.3))
        plan.append((ENSEMBLE_GRIDS[r % 4], r, T, sd))
    return plan


def make_contexts(mg, kws, threads=16):
    """Creates and initialises contexts on parallel host threads (the MT19937 initial
    state of each replica is drawn on the host in the reference's order)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda kw: mg.Stz(**kw), kws))


def count_events(recs):
    return sum(1 for r in recs if r.kind == 1)


def ensemble_leg(a, mg):
    """Batched mixed-precision ensemble vs the plain FP64 path on the same replicas."""
    plan = ensemble_plan(a.ensemble_replicas)
    kws = [mg.config_3d(n=g, seed=seed, q0_std=sd, temperature=T, precision=mg.MIXED) for g, seed, T, sd in plan]
    t0 = time.perf_counter()
    members = make_contexts(mg, kws)
    for m in members:
        m.set_print(5, 1.0)
        m.autotune(reps=5)          # measured once per grid shape, cached for the rest
    init_s = time.perf_counter() - t0
    print(f"# ensemble: {len(members)} replicas initialised in {init_s:.1f}s", file=sys.stderr, flush=True)
    # FP64 twins of one replica per grid, stepped alone, for the precision comparison
    twin_ids = list(range(min(4, len(plan))))
    twins = make_contexts(mg, [dict(kws[i], precision=mg.FP64) for i in twin_ids])
    for t in twins:
        t.set_print(5, 1.0)
        t.autotune(reps=5)
    ens = mg.Ensemble(members, batch=64)
    ens.run(a.ensemble_warmup, keep=False)
    for t in twins:
        t.run_steps(a.ensemble_warmup)
    t0 = time.perf_counter()
    rcs, outs = ens.run(a.ensemble_steps)
    dt = time.perf_counter() - t0
    events = [count_events(o) for o in outs]
    steps = [len(o) for o in outs]
    value = sum(events) / dt
    # plain FP64 path, one replica per grid: seconds per step and per-step agreement
    t_step = {}
    cmp = []
    for j, i in enumerate(twin_ids):
        t0 = time.perf_counter()
        _, recs64 = twins[j].run_steps(a.ensemble_steps)
        t_step[plan[i][0]] = (time.perf_counter() - t0) / max(len(recs64), 1)
        same = 0
        for x, y in zip(outs[i], recs64):
            if (x.kind, x.site, x.mode, x.n_transformed) != (y.kind, y.site, y.mode, y.n_transformed):
                break
            same += 1
        s64 = np.array([r.stress[5] for r in recs64[:same]])
        smx = np.array([r.stress[5] for r in outs[i][:same]])
        ss = float(np.max(np.abs(smx - s64)) / max(np.max(np.abs(s64)), 1e-30)) if same else None
        ferr = None
        if same == len(recs64) == len(outs[i]):
            f64 = twins[j].download(mg.F_SIGMA)
            fmx = members[i].download(mg.F_SIGMA)
            ferr = float(np.linalg.norm(fmx - f64) / np.linalg.norm(f64))
        cmp.append({"replica": i, "grid": list(plan[i][0]), "steps_identical": same, "ss_curve_rel_err": ss,
                    "stress_rel_err": ferr})
    fp64_time = sum(steps[i] * t_step[plan[i][0]] for i in range(len(plan)))
    fp64_value = sum(events) / fp64_time
    for t in twins:
        t.close()
    ens.close()
    for m in members:
        m.close()
    return {"value": round(value, 2), "unit": "replica-events/s", "replicas": len(plan),
            "grids": [list(g) for g in ENSEMBLE_GRIDS], "precision": "mixed", "steps_per_replica": a.ensemble_steps,
            "events": int(sum(events)), "seconds": round(dt, 3),
            "fp64_plain_equivalent": round(fp64_value, 2),
            "speedup_vs_fp64_plain": round(value / fp64_value, 3),
            "fp64_plain_ms_per_step": {"x".join(map(str, g)): round(v * 1e3, 4) for g, v in t_step.items()},
            "precision_check": cmp, "init_seconds": round(init_s, 1),
            "sample": "T and barrier std per replica from MT19937(4242); replica r uses grid r mod 4 and seed r"}


if __name__ == "__main__":
    main()
