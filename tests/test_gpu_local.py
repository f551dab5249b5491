"""GPU: the event-local stress update (stz_stress_update_local, resident SOLVER_LOCAL).

After transforms at a few sites, the field as last solved + C0:(mean-strain change) +
the shifted impulse responses of the sites must equal a full FFT solve of the new
eigenstrain field: within 1e-12 of the field's scale in FP64 (the full solve's own
rounding level; the hot-path contract vs the oracle is 1e-10), within 1e-5 in mixed
precision against the FP64 solve.  The resident step in SOLVER_LOCAL mode must produce
the FFT mode's event sequence (kinds, sites, modes, avalanche sizes) with mean stresses
within 1e-10, and replay the oracle's committed 10^4-event traces.
"""
import ast
import os

import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg
from trace_check import run_resident_with_snapshots

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _kw(dims, **extra):
    if len(dims) == 2:
        kw = dict(dim=2, n=(dims[0], dims[1], 1))
    else:
        kw = dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    kw.update(extra)
    return kw


def _random_fields(g, seed):
    rng = np.random.default_rng(seed)
    N, C, S, M = g.N, g.C, g.S, g.M
    return {
        mg.F_EPS0: rng.standard_normal(C * N) * 0.05,
        mg.F_Q0: 2.0 + 0.25 * rng.standard_normal(M * N),
        mg.F_DEPS: rng.standard_normal(M * S * N) * 0.07,
        mg.F_DEPS_LAST: rng.standard_normal(S * N) * 0.07,
        mg.F_PERM: rng.uniform(0, 0.5, N),
        mg.F_TEMP: rng.uniform(0, 2.0, N),
        mg.F_TLAST: -rng.uniform(0, 5e-7, N),
    }, rng.standard_normal(6) * 0.01


def _twin(kw, fields, epsavg):
    g = mg.Stz(init_state=False, **kw)
    for f, a in fields.items():
        g.upload(f, a)
    g.set_scalars(epsavg, np.zeros(6), 1e-8)
    g.stress_solve()
    return g


def _events(g, rng, nsites):
    """distinct sites (an avalanche transforms each voxel once), the grid's first and last
    voxel among them"""
    inner = rng.choice(np.arange(1, g.N - 1), size=max(nsites - 2, 0), replace=False)
    sites = ([0, g.N - 1] + [int(x) for x in inner])[:nsites]
    modes = [int(x) for x in rng.integers(g.M, size=len(sites))]
    redraws = [rng.standard_normal(g.M * g.S) * 0.07 for _ in sites]
    return sites, modes, redraws


def _drive(g, sites, modes, redraws, depsavg, ucomp, udelta):
    for s, m, r in zip(sites, modes, redraws):
        g.apply_event(s, m, r)
    e, st, t = g.scalars()
    g.set_scalars(e + depsavg, st, t)
    g.add_uniform_stress(ucomp, udelta)


CASES = [(32, 16), (512, 512), (16, 8, 32), (64, 64, 64), (96, 96, 96), (128, 64, 32)]


@pytest.mark.parametrize("dims", CASES)
@pytest.mark.parametrize("nsites", [1, 5, 64])
def test_update_local_equals_full_solve_fp64(dims, nsites):
    print("local fp64", dims, nsites, flush=True)
    kw = _kw(dims)
    probe = mg.Stz(init_state=False, **kw)
    fields, ea = _random_fields(probe, 3 + nsites)
    probe.close()
    full = _twin(kw, fields, ea)
    loc = _twin(kw, fields, ea)
    rng = np.random.default_rng(100 + nsites)
    sites, modes, redraws = _events(full, rng, nsites)
    dea = rng.standard_normal(6) * 1e-3
    uc = full.C - 1
    for g in (full, loc):
        _drive(g, sites, modes, redraws, dea, uc, 0.01)
    full.stress_solve()
    loc.stress_update_local(sites)
    a = full.download(mg.F_SIGMA)
    b = loc.download(mg.F_SIGMA)
    scale = np.max(np.abs(a))
    assert np.max(np.abs(a - b)) <= 1e-12 * scale
    sa, sb = full.scalars()[1], loc.scalars()[1]
    assert np.max(np.abs(sa - sb)) <= 1e-12 * scale
    # and against the oracle's solve of the same eigenstrain field and mean strain
    o = po.Oracle(empty=True, **kw)
    o.field(po.F_EPS0)[:] = full.download(mg.F_EPS0)
    e, _, _ = full.scalars()
    o.set_scalars(e, np.zeros(6), 0.0)
    o.solve()
    assert np.max(np.abs(o.field(po.F_SIGMA) - b)) <= 1e-10 * scale


@pytest.mark.parametrize("dims", [(32, 16), (16, 8, 32), (64, 64, 64)])
def test_update_local_mixed_against_fp64(dims):
    kw = _kw(dims)
    probe = mg.Stz(init_state=False, **kw)
    fields, ea = _random_fields(probe, 9)
    probe.close()
    full = _twin(kw, fields, ea)
    loc = _twin(dict(kw, precision=mg.MIXED), fields, ea)
    rng = np.random.default_rng(5)
    for _ in range(20):
        sites, modes, redraws = _events(full, rng, int(rng.integers(1, 9)))
        dea = rng.standard_normal(6) * 1e-4
        for g in (full, loc):
            _drive(g, sites, modes, redraws, dea, full.C - 1, 0.003)
        full.stress_solve()
        loc.stress_update_local(sites)
    a = full.download(mg.F_SIGMA)
    b = loc.download(mg.F_SIGMA)
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    print("mixed local vs fp64 full", dims, err, flush=True)
    assert err <= 1e-5


def test_update_local_long_sequence_fp64():
    """300 updates in a row stay at the full solve's rounding level."""
    kw = _kw((32, 32, 32))
    probe = mg.Stz(init_state=False, **kw)
    fields, ea = _random_fields(probe, 21)
    probe.close()
    full = _twin(kw, fields, ea)
    loc = _twin(kw, fields, ea)
    rng = np.random.default_rng(8)
    for i in range(300):
        sites, modes, redraws = _events(full, rng, 1 + (i % 3))
        dea = rng.standard_normal(6) * 1e-5
        for g in (full, loc):
            _drive(g, sites, modes, redraws, dea, 5, 1e-3)
        loc.stress_update_local(sites)
    full.stress_solve()
    a = full.download(mg.F_SIGMA)
    b = loc.download(mg.F_SIGMA)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("precision", [mg.FP64, mg.MIXED])
def test_zero_site_update_equals_resolve_after_a_run(precision):
    """After resident steps the field also holds the loading increments since its last
    solve; a zero-site update brings it to what a full solve of the current state gives."""
    kw = _kw3((16, 16, 16), precision=precision)
    g = mg.Stz(**kw)
    g.set_print(5, 1.0)
    g.set_solver(mg.SOLVER_LOCAL, 1024, 16)
    g.run_steps(300)
    twin = mg.Stz(init_state=False, **dict(kw, precision=mg.FP64))
    twin.upload(mg.F_EPS0, g.download(mg.F_EPS0))
    e, st, t = g.scalars()
    twin.set_scalars(e, st, t)
    twin.stress_solve()
    g.stress_update_local([])
    a, b = twin.download(mg.F_SIGMA), g.download(mg.F_SIGMA)
    err = np.linalg.norm(b - a) / np.linalg.norm(a)
    assert err <= (1e-12 if precision == mg.FP64 else 1e-5), err


def test_update_local_rejects_a_stale_field():
    kw = _kw((16, 8, 32))
    g = mg.Stz(init_state=False, **kw)
    g.stress_solve()
    g.upload(mg.F_EPS0, np.zeros(g.C * g.N))
    with pytest.raises(mg.StzError):
        g.stress_update_local([3])
    g.stress_solve()
    g.stress_update_local([3])
    with pytest.raises(mg.StzError):
        g.stress_update_local(list(range(65)))
    with pytest.raises(mg.StzError):
        g.stress_update_local([5, 9, 5])


def _kw3(n, **extra):
    kw = mg.config_3d(n=n, seed=1234, q0_std=0.2)
    kw.update(extra)
    return kw


def _run(kw, nsteps, mode, resync=64, local_max=16, printperiod=5):
    g = mg.Stz(**kw)
    g.set_print(printperiod, 1.0)
    g.set_solver(mode, resync, local_max)
    rc, recs = g.run_steps(nsteps)
    return g, rc, recs


RESIDENT = [
    ("3d16_shear", _kw3((16, 16, 16)), 1500),
    ("3d_16x8x32_shear", _kw3((16, 8, 32)), 1500),
    ("2d32_reference", dict(dim=2, n=(32, 32, 1), seed=123), 2000),
    ("2d32_shear", dict(dim=2, n=(32, 32, 1), seed=123, loading=1), 2000),
]


@pytest.mark.parametrize("name,kw,nsteps", RESIDENT, ids=[c[0] for c in RESIDENT])
def test_resident_local_equals_fft(name, kw, nsteps):
    print("resident local", name, flush=True)
    a, rca, ra = _run(kw, nsteps, mg.SOLVER_FFT)
    b, rcb, rb = _run(kw, nsteps, mg.SOLVER_LOCAL)
    assert rca == rcb and len(ra) == len(rb)
    lc = 0 if kw.get("loading", 0) == 0 else (2 if kw["dim"] == 2 else 5)
    scale = max(abs(r.stress[lc]) for r in ra) + 1e-30
    for i, (x, y) in enumerate(zip(ra, rb)):
        assert (x.kind, x.site, x.mode, x.n_transformed, x.n_negative, x.printed) == \
            (y.kind, y.site, y.mode, y.n_transformed, y.n_negative, y.printed), i
        assert np.max(np.abs(np.array(x.stress) - np.array(y.stress))) <= 1e-10 * scale, i
        assert abs(x.totime - y.totime) <= 1e-12 * abs(x.totime), i
    nth = sum(r.kind == mg.THERMAL for r in ra)
    nath = sum(r.kind == mg.ATHERMAL for r in ra)
    assert nth > 50
    fa = a.download(mg.F_SIGMA)
    fb = b.download(mg.F_SIGMA)
    assert np.max(np.abs(fa - fb)) <= 1e-10 * np.max(np.abs(fa))
    assert np.array_equal(a.download(mg.F_EPS0), b.download(mg.F_EPS0))
    print(f"{name}: {nth} thermal, {nath} athermal steps identical", flush=True)


def test_resident_local_mixed_tracks_fp64():
    kw = _kw3((32, 32, 32), precision=mg.MIXED)
    _, _, r64 = _run(dict(kw, precision=mg.FP64), 800, mg.SOLVER_FFT)
    _, _, rmx = _run(kw, 800, mg.SOLVER_LOCAL, resync=128)
    same = 0
    for x, y in zip(r64, rmx):
        if (x.kind, x.site, x.mode, x.n_transformed) != (y.kind, y.site, y.mode, y.n_transformed):
            break
        same += 1
    assert same > 100
    s64 = np.array([r.stress[5] for r in r64[:same]])
    smx = np.array([r.stress[5] for r in rmx[:same]])
    err = float(np.max(np.abs(smx - s64)) / np.max(np.abs(s64)))
    print("mixed local vs fp64 fft: identical steps", same, "curve err", err, flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("resync,local_max", [(1024, 16), (100, 1)], ids=["device", "host_fallbacks"])
@pytest.mark.parametrize("name", ["shear3d_64", "ref2d_512", "shear3d_128"])
def test_resident_local_replays_golden_trace(name, resync, local_max):
    """(100, 1): every 2-voxel avalanche and every 100th update run on the host path."""
    path = os.path.join(GOLD, f"trace_{name}.npz")
    if not os.path.exists(path):
        pytest.skip("fixture absent")
    d = np.load(path)
    kw = dict(ast.literal_eval(str(d["params"])))
    n = len(d["kind"])
    print("resident local trace", name, n, "steps", flush=True)
    g = mg.Stz(**kw)
    g.set_print(int(d["printperiod"]), float(d["endstrain"]))
    g.set_solver(mg.SOLVER_LOCAL, resync, local_max)
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
    print(f"{name}: {int(th.sum())} events identical with the local update, worst stress {worst:.2e}, "
          f"field snapshots {errs}", flush=True)


def test_ensemble_local_equals_single_runs():
    spec = [((16, 16, 16), 0, 300.0, 0.2), ((8, 16, 32), 1, 400.0, 0.15), ((16, 16, 16), 2, 260.0, 0.3)]
    members, refs = [], []
    for prec in (mg.FP64, mg.MIXED):
        for n, seed, T, sd in spec:
            kw = mg.config_3d(n=n, seed=seed, q0_std=sd, temperature=T, precision=prec)
            m = mg.Stz(**kw)
            m.set_print(5, 1.0)
            m.set_solver(mg.SOLVER_LOCAL, 32, 4)
            members.append(m)
            refs.append(_run(kw, 300, mg.SOLVER_LOCAL, resync=32, local_max=4))
    ens = mg.Ensemble(members, batch=32)
    rcs, outs = ens.run(300)
    for i, (ref, rc_ref, recs_ref) in enumerate(refs):
        assert rcs[i] == rc_ref and len(outs[i]) == len(recs_ref)
        for j, (a, b) in enumerate(zip(outs[i], recs_ref)):
            assert (a.kind, a.site, a.mode, a.n_transformed) == (b.kind, b.site, b.mode, b.n_transformed), (i, j)
            assert list(a.stress) == list(b.stress), (i, j)
        assert np.array_equal(members[i].download(mg.F_SIGMA), ref.download(mg.F_SIGMA))
    ens.close()
