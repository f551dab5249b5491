"""GPU: the local temperature (STZ_F_TLOCAL) and HeatConducting (stz_heat_conduct,
stz_set_heat_conduct) against the oracle, which the CPU tests pin with single Fourier
modes, a numpy restatement and hand-computed heating.

Tolerances: one diffusion 1e-12 of the field (FP64) / 1e-6 (mixed: FP32 spectrum);
trajectories with the flag on reproduce the oracle's event sequence with Tlocal within
1e-10 (FP64).
"""
import os

import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu


def _kw(dims, **extra):
    if len(dims) == 2:
        kw = dict(dim=2, n=(dims[0], dims[1], 1))
    else:
        kw = dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    kw.update(extra)
    return kw


CASES = [(16, 8), (64, 64), (512, 512), (8, 8, 8), (16, 8, 32), (64, 64, 64), (96, 96, 96)]


@pytest.mark.parametrize("precision", [mg.FP64, mg.MIXED])
@pytest.mark.parametrize("dims", CASES)
def test_heat_conduct_matches_oracle(dims, precision):
    print("heat", dims, precision, flush=True)
    kw = _kw(dims)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **dict(kw, precision=precision))
    assert np.all(g.download(mg.F_TLOCAL) == kw.get("temperature", 300.0))
    rng = np.random.default_rng(4)
    T = 300 + 20 * rng.standard_normal(o.N)
    o.field(po.F_TLOCAL)[:] = T
    g.upload(mg.F_TLOCAL, T)
    for dt in (0.05, 1.3):
        assert o.heat_conduct(dt) == 0
        g.heat_conduct(dt)
        ref = o.field(po.F_TLOCAL)
        err = np.max(np.abs(g.download(mg.F_TLOCAL) - ref)) / np.max(np.abs(ref))
        print("  dt", dt, "err %.2e" % err, flush=True)
        assert err < (1e-12 if precision == mg.FP64 else 1e-6)


def _pair(kw):
    o = po.Oracle(**kw)
    g = mg.Stz(**kw)
    return o, g


@pytest.mark.parametrize("heat", [False, True])
@pytest.mark.parametrize("dims,nsteps", [((32, 32), 600), ((8, 8, 8), 600)])
def test_trajectory_tlocal_matches_oracle(dims, nsteps, heat):
    """Transforms heat Tlocal (flag or not); with the flag every thermal and elastic step
    diffuses it.  Event sequence identical, Tlocal within 1e-10."""
    print("heat trajectory", dims, heat, flush=True)
    kw = _kw(dims, seed=11, loading=1)
    o, g = _pair(kw)
    o.set_heat(heat)
    g.set_heat_conduct(heat)
    o.set_print(5, 1.0)
    g.set_print(5, 1.0)
    events = 0
    for i in range(nsteps):
        rc_o, io = o.step()
        rc_g, ig = g.step()
        assert (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed), i
        if io.kind == 1:
            assert (io.site, io.mode) == (ig.site, ig.mode), i
        events += io.kind != 0
        assert rc_o == rc_g
        if rc_o:
            break
    ref = o.field(po.F_TLOCAL)
    got = g.download(mg.F_TLOCAL)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print("  events", events, "Tlocal err %.2e" % err, "range", float(ref.min()), float(ref.max()), flush=True)
    assert events > 20
    assert err < 1e-10
    assert ref.max() > 300.0


def test_run_steps_with_heat_equals_host_steps(tmp_path):
    kw = _kw((32, 32), seed=19, loading=1)
    g1 = mg.Stz(**kw)
    g2 = mg.Stz(**kw)
    for g in (g1, g2):
        g.set_heat_conduct(True)
        g.set_print(5, 1.0)
    seq1 = []
    for _ in range(300):
        rc, r = g1.step()
        seq1.append((r.kind, r.site, r.mode, r.n_transformed))
        if rc:
            break
    rc2, res2 = g2.run_steps(len(seq1))
    assert seq1 == [(r.kind, r.site, r.mode, r.n_transformed) for r in res2]
    assert np.array_equal(g1.download(mg.F_TLOCAL), g2.download(mg.F_TLOCAL))
    # a checkpoint keeps Tlocal and the switch
    path = os.path.join(str(tmp_path), "heat.ckpt")
    g1.checkpoint_save(path)
    a = [g1.step()[1] for _ in range(50)]
    h = mg.Stz(init_state=False, **kw)
    h.checkpoint_load(path)
    b = [h.step()[1] for _ in range(50)]
    assert [(r.kind, r.site, r.mode) for r in a] == [(r.kind, r.site, r.mode) for r in b]
    assert np.array_equal(g1.download(mg.F_TLOCAL), h.download(mg.F_TLOCAL))
    with pytest.raises(mg.StzError) as e:
        mg.Ensemble([g2])
    assert e.value.code == -5
