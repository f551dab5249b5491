"""GPU: the elastically inhomogeneous stress solve (stz_set_moduli + the fixed-point loop,
InhomogeneousStressSolver with per-voxel Lambda, Mu: STZ/elastic.c:977-1066) against the
oracle's restatement, which the CPU tests pin with uniform moduli and laminate closed
forms.

Tolerances: FP64 stress field 1e-10 of its largest value and the same number of
homogeneous solves; mixed precision 1e-4 and at most one solve more or fewer (the stop
test compares a change of about 1e-5 against its threshold).  Steps: the event sequence
exact and <sigma> within 1e-10 in FP64.
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


def _moduli(o, seed, contrast=0.25):
    """Random per-voxel moduli within +-contrast of the reference medium."""
    rng = np.random.default_rng(seed)
    lam0, mu0 = o.consts[3], o.consts[2]
    lam = lam0 * (1 + contrast * rng.uniform(-1, 1, o.N))
    mu = mu0 * (1 + contrast * rng.uniform(-1, 1, o.N))
    return lam, mu


def _state(o, g, seed):
    rng = np.random.default_rng(seed)
    eps0 = rng.standard_normal(o.C * o.N) * 0.03
    ea = rng.standard_normal(6) * 0.01
    o.field(po.F_EPS0)[:] = eps0
    o.set_scalars(ea, np.zeros(6), 0.0)
    g.upload(mg.F_EPS0, eps0)
    g.set_scalars(ea, np.zeros(6), 0.0)


@pytest.mark.parametrize("dims", [(16, 8), (8, 4, 16)])
def test_explicit_uniform_moduli_equal_the_uniform_solve(dims):
    """Lambda = lambda, Mu = mu everywhere: one homogeneous solve and, bit for bit, the
    stress of the uniform path (the reference's exit after its first iteration)."""
    print("inhom uniform", dims, flush=True)
    kw = _kw(dims)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **kw)
    _state(o, g, 3)
    g.stress_solve()
    s_uni = g.download(mg.F_SIGMA)
    m_uni = g.scalars()[1]
    g.set_moduli(np.full(o.N, o.consts[3]), np.full(o.N, o.consts[2]))
    g.stress_solve()
    assert g.inhom_iterations() == 1
    assert np.array_equal(g.download(mg.F_SIGMA), s_uni)
    assert np.array_equal(g.scalars()[1], m_uni)
    g.set_moduli(None, None)
    g.stress_solve()
    assert np.array_equal(g.download(mg.F_SIGMA), s_uni)


CASES = [(16, 8), (64, 64), (256, 128), (8, 8, 8), (16, 8, 32), (32, 32, 32), (64, 64, 64)]


@pytest.mark.parametrize("dims", CASES)
def test_inhom_solve_fp64_matches_oracle(dims):
    print("inhom fp64", dims, flush=True)
    kw = _kw(dims)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **kw)
    _state(o, g, 11)
    lam, mu = _moduli(o, 5)
    o.set_moduli(lam, mu)
    g.set_moduli(lam, mu)
    o.solve()
    g.stress_solve()
    it_o, it_g = o.inhom_iterations(), g.inhom_iterations()
    print("  iterations", it_o, it_g, flush=True)
    assert it_o == it_g and 2 <= it_o <= 100
    ref = o.field(po.F_SIGMA)
    got = g.download(mg.F_SIGMA)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref)) / scale
    print("  sigma err %.2e" % err, flush=True)
    assert err < 1e-10
    _, s_o, _ = o.scalars()
    _, s_g, _ = g.scalars()
    assert np.max(np.abs(s_g[: o.C] - s_o[: o.C])) < 1e-10 * scale


@pytest.mark.parametrize("dims", [(64, 64), (32, 32, 32), (64, 64, 64)])
def test_inhom_solve_mixed_matches_oracle(dims):
    print("inhom mixed", dims, flush=True)
    kw = _kw(dims)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **dict(kw, precision=mg.MIXED))
    _state(o, g, 12)
    lam, mu = _moduli(o, 6)
    o.set_moduli(lam, mu)
    g.set_moduli(lam, mu)
    o.solve()
    g.stress_solve()
    assert abs(o.inhom_iterations() - g.inhom_iterations()) <= 1
    ref = o.field(po.F_SIGMA)
    got = g.download(mg.F_SIGMA)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print("  sigma err %.2e" % err, flush=True)
    assert err < 1e-4


def test_strong_contrast_stops_at_the_cap():
    kw = _kw((32, 32))
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **kw)
    _state(o, g, 2)
    fac = np.where(np.arange(o.N) % 2 == 0, 1.0, 9.0)
    lam, mu = o.consts[3] * fac, o.consts[2] * fac
    o.set_moduli(lam, mu)
    g.set_moduli(lam, mu)
    o.solve()
    g.stress_solve()
    assert o.inhom_iterations() == 101 and g.inhom_iterations() == 101
    ref = o.field(po.F_SIGMA)
    got = g.download(mg.F_SIGMA)
    assert np.max(np.abs(got - ref)) < 1e-8 * np.max(np.abs(ref))


@pytest.mark.parametrize("dims", [(16, 8), (8, 8, 8)])
def test_inhom_displacement_matches_oracle(dims):
    print("inhom displacement", dims, flush=True)
    kw = _kw(dims)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, **kw)
    _state(o, g, 21)
    lam, mu = _moduli(o, 8)
    o.set_moduli(lam, mu)
    g.set_moduli(lam, mu)
    uo = o.displacement()
    ug = g.displacement()
    err = np.max(np.abs(ug - uo)) / np.max(np.abs(uo))
    assert err < 1e-10, err
    # the call re-solved the stress, as the reference's uflag=2 call does
    ref = o.field(po.F_SIGMA)
    assert np.max(np.abs(g.download(mg.F_SIGMA) - ref)) < 1e-10 * np.max(np.abs(ref))


def _pair_steps(kw, seed):
    o = po.Oracle(**kw)
    g = mg.Stz(**kw)
    lam, mu = _moduli(o, seed)
    o.set_moduli(lam, mu)
    g.set_moduli(lam, mu)
    o.solve()               # the initial solve with the moduli (STZ/bmg.c:117-127)
    g.stress_solve()
    return o, g, lam, mu


@pytest.mark.parametrize("dims,nsteps", [((32, 32), 1500), ((8, 8, 8), 600)])
def test_inhom_trajectory_matches_oracle(dims, nsteps):
    print("inhom trajectory", dims, flush=True)
    kw = _kw(dims, seed=321, loading=1)
    o, g, _, _ = _pair_steps(kw, 4)
    o.set_print(5, 1.0)
    g.set_print(5, 1.0)
    worst, events, iters = 0.0, 0, []
    for i in range(nsteps):
        rc_o, io = o.step()
        rc_g, ig = g.step()
        assert (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed), i
        if io.kind != 0:
            events += 1
        if io.kind == 1:
            assert (io.site, io.mode) == (ig.site, ig.mode), i
        so = np.array(io.stress[:6])
        sg = np.array(ig.stress[:6])
        worst = max(worst, np.max(np.abs(so - sg)) / max(np.max(np.abs(so)), 1e-3))
        iters.append(g.inhom_iterations())
        assert rc_o == rc_g
        if rc_o:
            break
    print("  events", events, "worst %.2e" % worst, "max iterations", max(iters), flush=True)
    assert events > 60                  # thermal events and avalanches
    assert worst < 1e-10
    assert max(iters) > 1


def test_inhom_run_steps_equals_host_steps():
    """stz_run_steps takes the host path with per-voxel moduli: same steps as stz_step."""
    kw = _kw((32, 32), seed=77, loading=1)
    o, g1, lam, mu = _pair_steps(kw, 9)
    g2 = mg.Stz(**kw)
    g2.set_moduli(lam, mu)
    g2.stress_solve()
    for g in (g1, g2):
        g.set_print(5, 1.0)
    seq1 = []
    for _ in range(120):
        rc, r = g1.step()
        seq1.append((r.kind, r.site, r.mode, r.n_transformed, tuple(r.stress[:3])))
        if rc:
            break
    rc2, res2 = g2.run_steps(len(seq1))
    seq2 = [(r.kind, r.site, r.mode, r.n_transformed, tuple(r.stress[:3])) for r in res2]
    assert seq1 == seq2
    assert np.array_equal(g1.download(mg.F_SIGMA), g2.download(mg.F_SIGMA))


def test_inhom_refuses_local_update_and_ensembles():
    kw = _kw((16, 16))
    g = mg.Stz(**kw)
    o = po.Oracle(empty=True, **kw)
    lam, mu = _moduli(o, 1)
    g.set_moduli(lam, mu)
    with pytest.raises(mg.StzError) as e:
        g.set_solver(mg.SOLVER_LOCAL)
    assert e.value.code == -5
    with pytest.raises(mg.StzError) as e:
        mg.Ensemble([g])
    assert e.value.code == -5
    g.set_moduli(None, None)
    g.set_solver(mg.SOLVER_LOCAL)
    with pytest.raises(mg.StzError) as e:
        g.set_moduli(lam, mu)
    assert e.value.code == -5


def test_inhom_checkpoint_restart(tmp_path):
    kw = _kw((32, 32), seed=55, loading=1)
    o, g, lam, mu = _pair_steps(kw, 3)
    g.set_print(5, 1.0)
    for _ in range(40):
        g.step()
    path = os.path.join(str(tmp_path), "inhom.ckpt")
    g.checkpoint_save(path)
    a = [g.step()[1] for _ in range(60)]
    h = mg.Stz(init_state=False, **kw)
    with pytest.raises(mg.StzError):
        h.checkpoint_load(path)            # uniform moduli: refused
    h.set_moduli(lam, mu)
    h.checkpoint_load(path)
    b = [h.step()[1] for _ in range(60)]
    assert [(r.kind, r.site, r.mode) for r in a] == [(r.kind, r.site, r.mode) for r in b]
    assert np.array_equal(g.download(mg.F_SIGMA), h.download(mg.F_SIGMA))


def test_lame_mirrors_the_library_constants():
    """Stz.lame() gives the reference medium's constants as the oracle derives them."""
    for kw in (_kw((16, 8)), _kw((8, 8, 8)), _kw((16, 8), double_nu=0)):
        g = mg.Stz(init_state=False, **kw)
        o = po.Oracle(empty=True, **kw)
        assert g.lame() == (o.consts[3], o.consts[2])
