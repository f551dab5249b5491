"""GPU: the device-resident step (stz_run_steps) and the batched ensemble.

The resident step must produce, step for step, the same records as the per-call host
loop (stz_step) from the same state: identical kinds, events, avalanche sizes and
bit-identical clock, mean strain and mean stress (both paths run the same kernels in
the same order and consume the same MT19937 stream).  It is checked in graph replay
and in direct launches, with batch sizes and lookup windows small enough that halts
(print steps, window refills, step budgets) happen inside the run, and against the
oracle's committed 10^4-event traces.  Each ensemble replica must equal the same
context run alone.
"""
import ast
import os

import numpy as np
import pytest

import metallic_glasses_amd as mg
from trace_check import run_resident_with_snapshots

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _kw3(n, **extra):
    kw = mg.config_3d(n=n, seed=1234, q0_std=0.2)
    kw.update(extra)
    return kw


def _same_record(a, b, i, soc_rtol=1e-12):
    assert (a.kind, a.site, a.mode, a.n_transformed, a.n_negative, a.printed) == \
        (b.kind, b.site, b.mode, b.n_transformed, b.n_negative, b.printed), i
    assert a.q_chosen == b.q_chosen and a.total_rate == b.total_rate, i
    assert a.totime == b.totime, i
    assert list(a.epsavg) == list(b.epsavg), i
    assert list(a.stress) == list(b.stress), i
    assert abs(a.soc_elast - b.soc_elast) <= soc_rtol * max(abs(a.soc_elast), 1e-300), i


def _stepwise(kw, nsteps, printperiod=5, endstrain=1.0):
    g = mg.Stz(**kw)
    g.set_print(printperiod, endstrain)
    out = []
    rc = 0
    for _ in range(nsteps):
        rc, r = g.step()
        out.append(r)
        if rc:
            break
    return g, rc, out


def _compare_fields(a, b):
    for f in (mg.F_EPS0, mg.F_SIGMA, mg.F_DEPS, mg.F_DEPS_LAST, mg.F_PERM, mg.F_TEMP, mg.F_TLAST, mg.F_Q0):
        assert np.array_equal(a.download(f), b.download(f)), f
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert np.array_equal(a.counters(), b.counters())


CASES = [
    # name, params, steps, batch, graph, window words
    ("3d16_graph", _kw3((16, 16, 16)), 600, 64, 1, 0),
    ("3d16_graph8", _kw3((16, 16, 16)), 600, 60, 8, 0),
    ("3d16_direct_small_batches", _kw3((16, 16, 16)), 600, 7, 0, 0),
    ("3d16_small_window", _kw3((16, 16, 16)), 600, 64, 0, 8192),
    ("3d_16x8x32_mixed", _kw3((16, 8, 32), precision=mg.MIXED), 500, 32, 0, 0),
    ("2d32_reference", dict(dim=2, n=(32, 32, 1), seed=123, loading=1), 1500, 64, 2, 0),
]


@pytest.mark.parametrize("name,kw,nsteps,batch,graph,window", CASES, ids=[c[0] for c in CASES])
def test_resident_equals_stepwise(name, kw, nsteps, batch, graph, window):
    print("resident", name, flush=True)
    ref, rc_ref, recs_ref = _stepwise(kw, nsteps)
    g = mg.Stz(**kw)
    g.set_print(5, 1.0)
    g.run_config(batch=batch, use_graph=graph, window_words=window)
    rc, recs = g.run_steps(nsteps)
    assert rc == rc_ref
    assert len(recs) == len(recs_ref)
    for i, (a, b) in enumerate(zip(recs, recs_ref)):
        _same_record(a, b, i)
    assert sum(r.kind == mg.THERMAL for r in recs) > 50
    _compare_fields(g, ref)
    # the host loop continues where the resident run stopped
    for i in range(20):
        ra, a = g.step()
        rb, b = ref.step()
        _same_record(a, b, i)


def test_resident_in_pieces_equals_one_run():
    kw = _kw3((16, 16, 16))
    a = mg.Stz(**kw)
    a.set_print(5, 1.0)
    _, ra = a.run_steps(400)
    b = mg.Stz(**kw)
    b.set_print(5, 1.0)
    rb = []
    for n in (1, 13, 86, 100, 200):
        _, part = b.run_steps(n)
        rb += part
    assert len(ra) == len(rb) == 400
    for i, (x, y) in enumerate(zip(ra, rb)):
        _same_record(x, y, i)


def test_resident_termination_and_stop_at_print():
    kw = _kw3((16, 16, 16))
    ref, rc_ref, recs_ref = _stepwise(kw, 5000, printperiod=2, endstrain=0.004)
    assert rc_ref == 1
    g = mg.Stz(**kw)
    g.set_print(2, 0.004)
    rc, recs = g.run_steps(5000)
    assert rc == 1 and len(recs) == len(recs_ref)
    for i, (a, b) in enumerate(zip(recs, recs_ref)):
        _same_record(a, b, i)
    # stop at print: every call returns right after a printed step
    h = mg.Stz(**kw)
    h.set_print(2, 0.004)
    h.stop_at_print(True)
    got = []
    for _ in range(400):
        rc, part = h.run_steps(5000)
        got += part
        if rc == 1:
            break
        assert part and part[-1].printed == 1
        last = part[-1]
        rc = h.step_print(last)        # the print-step re-solve, then the termination test
        got[-1] = last
        if rc == 1:
            break
    assert len(got) == len(recs_ref)
    for i, (a, b) in enumerate(zip(got, recs_ref)):
        _same_record(a, b, i)


@pytest.mark.parametrize("name", ["shear3d_64", "ref2d_512", "shear3d_128"])
def test_resident_replays_golden_trace(name):
    path = os.path.join(GOLD, f"trace_{name}.npz")
    if not os.path.exists(path):
        pytest.skip("fixture absent")
    d = np.load(path)
    kw = dict(ast.literal_eval(str(d["params"])))
    n = len(d["kind"])
    print("resident trace", name, n, "steps", flush=True)
    g = mg.Stz(**kw)
    g.set_print(int(d["printperiod"]), float(d["endstrain"]))
    recs, errs = run_resident_with_snapshots(g, d, n)
    assert len(recs) == n
    lc = 0 if kw.get("loading", 0) == 0 else (2 if kw["dim"] == 2 else 5)
    kinds = np.array([r.kind for r in recs])
    assert np.array_equal(kinds, d["kind"])
    th = kinds == 1
    assert np.array_equal(np.array([r.site for r in recs])[th], d["site"][th])
    assert np.array_equal(np.array([r.mode for r in recs])[th], d["mode"][th])
    ath = kinds == 2
    assert np.array_equal(np.array([r.n_transformed for r in recs])[ath], d["ntrans"][ath])
    st = np.array([r.stress[lc] for r in recs])
    worst = float(np.max(np.abs(st - d["stress"]) / np.maximum(np.abs(d["stress"]), 1e-3)))
    assert worst < 1e-10, worst
    tt = np.array([r.totime for r in recs])
    assert np.max(np.abs(tt - d["totime"]) / np.maximum(np.abs(d["totime"]), 1e-30)) <= 1e-10
    print(f"{name}: {int(th.sum())} events identical through the resident step, worst stress {worst:.2e}, "
          f"field snapshots {errs}", flush=True)


ENSEMBLE = [
    ((16, 16, 16), 0, 300.0, 0.2),
    ((8, 16, 32), 1, 400.0, 0.15),
    ((16, 16, 16), 2, 260.0, 0.3),
    ((32, 32, 32), 3, 350.0, 0.25),
    ((16, 16, 16), 4, 440.0, 0.1),
]


@pytest.mark.parametrize("precision", [mg.FP64, mg.MIXED])
def test_ensemble_replicas_equal_single_runs(precision):
    nsteps = 300
    members, refs = [], []
    for n, seed, T, sd in ENSEMBLE:
        kw = mg.config_3d(n=n, seed=seed, q0_std=sd, temperature=T, precision=precision)
        m = mg.Stz(**kw)
        m.set_print(5, 1.0)
        members.append(m)
        refs.append(_stepwise(kw, nsteps))
    ens = mg.Ensemble(members, batch=32, use_graph=4 if precision == mg.MIXED else 0)
    rcs, outs = ens.run(nsteps)
    for i, (ref, rc_ref, recs_ref) in enumerate(refs):
        print("replica", i, ENSEMBLE[i], flush=True)
        assert rcs[i] == rc_ref
        assert len(outs[i]) == len(recs_ref)
        for j, (a, b) in enumerate(zip(outs[i], recs_ref)):
            _same_record(a, b, (i, j))
        _compare_fields(members[i], ref)
    ens.close()
