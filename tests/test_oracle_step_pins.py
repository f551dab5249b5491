"""Hand-computed pins of the oracle's step logic: ActEnergy's work and direction terms,
the athermal avalanche, the thermal and elastic loading branches with the clock, the print
schedule with the barrier reset, and both termination tests.

Every expected value comes from a closed form or from mathematics the oracle does not
use (a matrix inverse of the Voigt stiffness, the mean of a periodic eigenstrain
problem, mixed strain/stress boundary conditions), so a dropped term, a wrong sign,
component or index fails a check.  Reference lines:
ActEnergy STZ/init.c:561-690, athermal_plasticity STZ/evolution.c:47-100, KMC loading and
clock STZ/evolution.c:262-358, print block and termination STZ/bmg.c:181-410.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from test_oracle_rng import _polar_from_uniforms

E_J = 1.60217733e-19
KB = 1.380658e-23
V = 0.7 * 0.7 * 1.0
THERMAL, ATHERMAL, ELASTIC = 1, 2, 0


def _stiffness(lam, mu, dim):
    """Voigt stiffness acting on tensor components: 2-D (xx, yy, xy), 3-D (xx, yy, zz, yz,
    xz, xy); shear rows give sigma_ij = 2 mu eps_ij."""
    if dim == 2:
        return np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, 2 * mu]])
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = lam
    for i in range(3):
        Cm[i, i] += 2 * mu
        Cm[3 + i, 3 + i] = 2 * mu
    return Cm


def _moduli(o):
    c = o.consts
    return dict(E=c[0], nu=c[1], mu=c[2], lam=c[3], kT=c[4], dt=c[6])


def _full(d, dim):
    """Independent components (2-D xx, xy; 3-D xx, yy, yz, xz, xy) -> tensor components."""
    d = np.asarray(d, dtype=float)
    if dim == 2:
        return np.array([d[0], -d[0], d[1]])
    return np.array([d[0], d[1], -(d[0] + d[1]), d[2], d[3], d[4]])


def _wdot(a, b, dim):
    w = np.array([1, 1, 2.0]) if dim == 2 else np.array([1, 1, 1, 2, 2, 2.0])
    return float(np.sum(w * a * b))


def _stream(seed, n=4000):
    return iter(po.mt_stream(seed, n).tolist())


# ---------------------------------------------------------------- ActEnergy
@pytest.mark.parametrize("dim", [2, 3])
def test_act_energy_work_and_direction_closed_form(dim):
    """Pure shear sigma parallel to the mode strain gives W = s |deps|_w^2 V 1e-18/e; a last
    strain parallel to the mode gives D = 1, antiparallel D = 0, orthogonal D = 1/4."""
    kw = dict(dim=dim, n=(2, 2, 1) if dim == 2 else (2, 2, 2), modes=3, softening=1)
    if dim == 3:
        kw.update(E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    o = po.Oracle(empty=True, **kw)
    N, C, S, M = o.N, o.C, o.S, o.M
    d0 = np.array([0.031, -0.052]) if dim == 2 else np.array([0.031, -0.012, 0.02, -0.044, 0.027])
    f0 = _full(d0, dim)
    # an independent-component vector whose tensor is w-orthogonal to f0
    if dim == 2:
        d1 = np.array([d0[1], -d0[0]])
    else:
        # remove from a fixed vector its component along f0 in the weighted inner product;
        # _full is linear, so <full(x), f0>_w = x . g with g_i = <full(e_i), f0>_w
        g = np.array([_wdot(_full(b, dim), f0, dim) for b in np.eye(5)])
        cand = np.array([0.02, 0.035, -0.01, 0.015, 0.04])
        d1 = cand - g * (cand @ g) / (g @ g)
    f1 = _full(d1, dim)
    assert abs(_wdot(f1, f0, dim)) < 1e-15
    deps = o.field(po.F_DEPS).reshape(M, S, N)
    for v in range(N):
        deps[0, :, v] = d0
        deps[1, :, v] = d1
        deps[2, :, v] = 0.5 * d0
    last = o.field(po.F_DEPS_LAST).reshape(S, N)
    # voxel 0: parallel last strain, voxel 1: antiparallel, voxel 2: none, others: parallel x3
    last[:, 0] = 0.4 * d0
    last[:, 1] = -0.7 * d0
    last[:, 2] = 0.0
    for v in range(3, N):
        last[:, v] = 3.0 * d0
    q0 = np.full(M * N, 1.7)
    o.set_barriers(q0)
    s_amp = 1.5
    sig = o.field(po.F_SIGMA).reshape(C, N)
    for v in range(N):
        sig[:, v] = s_amp * f0
    o.field(po.F_PERM)[:] = 0.4
    o.field(po.F_TEMP)[:] = 0.0
    o.softening()
    o.act_energy()
    q = o.field(po.F_Q).reshape(M, N)
    W0 = s_amp * _wdot(f0, f0, dim) * V * 1e-18 / E_J
    assert W0 > 0.01
    Sv = 0.4
    # mode 0, parallel last strain: D = 1
    assert abs(q[0, 0] - (1.7 * math.exp(-Sv) - W0 / 2)) < 1e-14
    # mode 0, antiparallel: D = 0
    assert abs(q[0, 1] - (1.7 - W0 / 2)) < 1e-14
    # no last strain: D = 1/4 (cosine 0)
    assert abs(q[0, 2] - (1.7 * math.exp(-Sv / 4) - W0 / 2)) < 1e-14
    # mode 1 is orthogonal to the stress and to the last strain: W = 0, D = 1/4
    for v in (0, 1, 3):
        assert abs(q[1, v] - 1.7 * math.exp(-Sv / 4)) < 1e-14
    # mode 2 = half of mode 0: half the work, same direction (D independent of magnitude)
    assert abs(q[2, 0] - (1.7 * math.exp(-Sv) - W0 / 4)) < 1e-14
    assert abs(q[2, 1] - (1.7 - W0 / 4)) < 1e-14


# ---------------------------------------------------------------- athermal
def test_athermal_hand_case_2x2():
    """2x2 grid, M = 2, softening off and sigma = 0, so Q = Q0.  Voxel 0 has a tie at a
    negative barrier (first mode wins), voxel 1 a minimum just inside 0.33, voxel 2 exactly
    at 0.33 (not transformed), voxel 3 a positive tie inside 0.33."""
    o = po.Oracle(empty=True, dim=2, n=(2, 2, 1), modes=2, softening=0, seed=77, athermal_correction=0)
    N, S, M = o.N, o.S, o.M
    k = _moduli(o)
    q0 = np.array([[-0.1, 0.5, 0.33, 0.2],      # mode 0 per voxel
                   [-0.1, 0.3299, 0.4, 0.2]])   # mode 1
    o.set_barriers(q0.ravel())
    deps = o.field(po.F_DEPS).reshape(M, S, N)
    base = np.arange(M * S * N, dtype=float).reshape(M, S, N) * 0.003 + 0.01
    deps[:] = base
    eps_bar = np.array([1e-3, -2e-4, 5e-4])
    o.set_scalars(np.r_[eps_bar, 0, 0, 0], np.zeros(6), 2.5)
    rc, info = o.step()
    assert rc == 0
    assert info.kind == ATHERMAL
    assert info.n_transformed == 3
    assert info.totime == 2.5                       # no clock advance on an athermal step
    eps0 = o.field(po.F_EPS0).reshape(3, N)
    picks = {0: 0, 1: 1, 3: 0}
    for v in range(N):
        if v in picks:
            m = picks[v]
            assert np.array_equal(eps0[:, v], _full(base[m, :, v], 2))
            assert np.array_equal(o.field(po.F_DEPS_LAST).reshape(S, N)[:, v], base[m, :, v])
            assert o.field(po.F_TLAST)[v] == 2.5
        else:
            assert np.all(eps0[:, v] == 0)
            assert np.array_equal(deps[:, :, v], base[:, :, v])
    # redraws in row-major order of the transformed voxels: M*S Gaussians each, mode-major
    it = _stream(77)
    for v in (0, 1, 3):
        g = [_polar_from_uniforms(it) for _ in range(M * S)]
        assert np.array_equal(deps[:, :, v].ravel(), 0.5 * 0.14 * np.array(g))
    # the solve's mean stress is C0 : (eps_bar - <eps0>) (the fluctuation has zero mean)
    C0 = _stiffness(k["lam"], k["mu"], 2)
    mean_s = C0 @ (eps_bar - eps0.mean(axis=1))
    assert np.allclose(np.array(info.stress[:3]), mean_s, rtol=0, atol=1e-13 * np.abs(mean_s).max())
    # eps_bar -= C0^-1 (0, <s_yy>, <s_xy>), the reference's subtraction (STZ/evolution.c:91-97)
    corr = np.linalg.inv(C0) @ np.array([0.0, info.stress[1], info.stress[2]])
    assert np.allclose(np.array(info.epsavg[:3]), eps_bar - corr, rtol=0, atol=1e-16)
    c = o.counters()
    assert (c[0], c[1], c[2], c[3]) == (1, 1, 0, 0)      # step, astep, kstep, estep


# ---------------------------------------------------------------- thermal / elastic
CASES = [
    ("2d_tension", dict(dim=2, n=(4, 4, 1), loading=0)),
    ("2d_shear", dict(dim=2, n=(4, 4, 1), loading=1)),
    ("3d_tension", dict(dim=3, n=(4, 2, 2), loading=0, E=94.5, nu=0.35, double_nu=0, gamma0=0.1,
                        athermal_correction=1)),
    ("3d_shear", dict(dim=3, n=(4, 2, 2), loading=1, E=94.5, nu=0.35, double_nu=0, gamma0=0.1,
                      athermal_correction=1)),
]


def _uniform_state(kw, q0, seed=5, **extra):
    o = po.Oracle(empty=True, modes=2, softening=0, seed=seed, **dict(kw, **extra))
    N, S, M = o.N, o.S, o.M
    rng = np.random.default_rng(11)
    o.field(po.F_DEPS)[:] = 0.05 * rng.standard_normal(M * S * N)
    o.set_barriers(np.full(M * N, q0))
    return o


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_thermal_step_clock_selection_and_mixed_loading(name, kw):
    """Equal rates: R = M N V nu0 exp(-Q/kT), t = 1/R, the dice picks index ceil(u M N) - 1.
    After the step the loaded strain grew by strainrate * t, and the mean stress of the new
    state, C0 : (eps_bar - <eps0>), equals the tracked <sigma> in the loaded component and is
    zero in all others (strain control along the load, zero mean stress across it)."""
    q = 0.6
    o = _uniform_state(kw, q, seed=9)
    N, C, S, M = o.N, o.C, o.S, o.M
    k = _moduli(o)
    dim = kw["dim"]
    lc = int(o.consts[15])
    eps_bar = np.zeros(6)
    eps_bar[lc] = 2e-3
    C0 = _stiffness(k["lam"], k["mu"], dim)
    s0 = np.zeros(6)
    s0[:C] = C0 @ eps_bar[:C]
    o.set_scalars(eps_bar, s0, 1.0)
    # the stored field stays zero, so the work term vanishes and all M*N rates are equal;
    # the event's solve recomputes the field from eps_bar and eps0
    deps_before = o.field(po.F_DEPS).reshape(M, S, N).copy()
    rc, info = o.step()
    assert rc == 0 and info.kind == THERMAL
    R = M * N * (V * 1e13) * math.exp(-q / k["kT"])
    t = 1 / R
    assert abs(info.total_rate - R) < 1e-13 * R
    assert abs(info.totime - (1.0 + t)) <= 2 * np.spacing(1.0)
    u = next(_stream(9)) / 4294967296.0
    ev = math.ceil(u * M * N) - 1
    assert (info.mode, info.site) == (ev // N, ev % N)
    # the transformed voxel carries the chosen mode's strain; all others none
    eps0 = o.field(po.F_EPS0).reshape(C, N)
    assert np.array_equal(eps0[:, info.site], _full(deps_before[info.mode, :, info.site], dim))
    assert np.count_nonzero(np.any(eps0 != 0, axis=0)) == 1
    e_new = np.array(info.epsavg[:C])
    sr = 1e-4
    # the solve's mean stress plus the load increment: E sr t in tension, 2 mu sr t in shear
    mean_solve = C0 @ (eps_bar[:C] - eps0.mean(axis=1))
    K = 2 * k["mu"] if kw["loading"] == 1 else k["E"]
    assert abs(info.stress[lc] - (mean_solve[lc] + K * sr * t)) < 1e-13 * abs(info.stress[lc])
    if kw["loading"] == 1:
        # shear components decouple, so the loaded strain grows by exactly strainrate * t
        assert abs(e_new[lc] - (eps_bar[lc] + sr * t)) < 1e-18
    mean_new = C0 @ (e_new - eps0.mean(axis=1))
    scale = abs(info.stress[lc])
    assert abs(mean_new[lc] - info.stress[lc]) < 1e-12 * scale
    others = [c for c in range(C) if c != lc]
    assert np.max(np.abs(mean_new[others])) < 1e-12 * scale
    # the stored field is the solve of the old mean strain plus the uniform load increment
    # on the loaded component; a re-solve of the new state differs from it by exactly the
    # old solve's mean on the other components (uniformly) and not at all on the loaded one
    field = o.field(po.F_SIGMA).reshape(C, N).copy()
    o.solve()
    diff = field - o.field(po.F_SIGMA).reshape(C, N)
    assert np.max(np.abs(diff[lc])) < 1e-12 * scale
    for c in others:
        assert np.ptp(diff[c]) < 1e-12 * scale
        assert abs(diff[c].mean() - info.stress[c]) < 1e-12 * scale
    c = o.counters()
    assert (c[0], c[1], c[2], c[3]) == (1, 0, 1, 0)


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_elastic_step_closed_form(name, kw):
    """1/R >= dt: the clock advances by dt = increment / strainrate and the load by one
    increment.  The strain rate (2e-4) differs from the increment (1e-4) so that the
    reference's shear branch, which adds strainrate*1 to the strain but 2 mu * increment to
    the stress (STZ/evolution.c:349-356), is told apart from a consistent one."""
    sr, inc = 2e-4, 1e-4
    o = _uniform_state(kw, 3.0, strainrate=sr, strainincrement=inc)
    N, C = o.N, o.C
    k = _moduli(o)
    dim = kw["dim"]
    lc = int(o.consts[15])
    assert k["dt"] == inc / sr
    e0 = np.zeros(6)
    e0[lc] = 1e-3
    s0 = np.zeros(6)
    C0 = _stiffness(k["lam"], k["mu"], dim)
    s0[:C] = C0 @ e0[:C]
    o.set_scalars(e0, s0, 4.0)
    o.field(po.F_SIGMA).reshape(C, N)[:] = s0[:C, None]
    mt_before = np.array(o.rng().contents.mt[:]), o.rng().contents.mti
    rc, info = o.step()
    assert rc == 0 and info.kind == ELASTIC
    assert info.totime == 4.0 + inc / sr
    de = np.array(info.epsavg[:C]) - e0[:C]
    ds = np.array(info.stress[:C]) - s0[:C]
    field = o.field(po.F_SIGMA).reshape(C, N)
    if kw["loading"] == 1:
        expect_de = np.zeros(C)
        expect_de[lc] = sr * 1.0
        expect_ds = np.zeros(C)
        expect_ds[lc] = 2 * k["mu"] * inc
    else:
        nu = k["nu"]
        expect_de = np.zeros(C)
        expect_de[0] = inc
        expect_de[1] = -nu * inc
        if dim == 3:
            expect_de[2] = -nu * inc
        expect_ds = np.zeros(C)
        expect_ds[0] = k["E"] * inc
        # uniaxial stress: the strain increment's stress is E*inc along x and zero across
        assert np.allclose(C0 @ expect_de, expect_ds, rtol=0, atol=1e-12 * k["E"] * inc)
    assert np.allclose(de, expect_de, rtol=0, atol=1e-19)
    assert np.allclose(ds, expect_ds, rtol=0, atol=1e-15)
    assert np.allclose(field - s0[:C, None], expect_ds[:, None], rtol=0, atol=1e-15)
    # no random number is consumed
    assert np.array_equal(np.array(o.rng().contents.mt[:]), mt_before[0])
    assert o.rng().contents.mti == mt_before[1]
    c = o.counters()
    assert (c[0], c[1], c[2], c[3]) == (1, 0, 0, 1)


# ---------------------------------------------------------------- print block / termination
def _elastic_run(kw, printperiod, endstrain, nsteps, stress0=None, print_on=True):
    o = _uniform_state(kw, 3.0)
    if stress0 is not None:
        s = np.zeros(6)
        s[int(o.consts[15])] = stress0
        o.set_scalars(np.zeros(6), s, 0.0)
    if print_on:
        o.set_print(printperiod, endstrain)
    out = []
    for _ in range(nsteps):
        rc, info = o.step()
        assert rc >= 0
        out.append((rc, info.printed))
        if rc == 1:
            break
    return o, out


def test_print_schedule_every_period_and_end_strain():
    kw = CASES[0][1]                                  # 2-D tension, increment 1e-4 per step
    _, out = _elastic_run(kw, 5, 0.00105, 50)
    printed = [i + 1 for i, (_, p) in enumerate(out) if p]
    assert printed == [5, 10]
    # |eps_xx| = 11e-4 > 0.00105 at step 11, the first step past the end strain
    assert len(out) == 11 and out[-1][0] == 1
    _, out = _elastic_run(kw, 3, 1.0, 20)
    assert [i + 1 for i, (_, p) in enumerate(out) if p] == [3, 6, 9, 12, 15, 18]


@pytest.mark.parametrize("period", [1, 0, -2])
def test_print_every_step_for_period_at_most_one(period):
    # |d strain| >= period*inc*0.99 holds at every elastic step for period <= 1
    _, out = _elastic_run(CASES[0][1], period, 1.0, 12)
    assert all(p == 1 for _, p in out) and len(out) == 12


def test_no_print_or_termination_before_set_print():
    _, out = _elastic_run(CASES[0][1], 5, 0.0, 15, stress0=-10.0, print_on=False)
    assert len(out) == 15 and all(rc == 0 and p == 0 for rc, p in out)


def test_tension_instability_terminates():
    # <sigma_xx> < -0.5 after the step (STZ/bmg.c:388)
    _, out = _elastic_run(CASES[0][1], 1000, 1.0, 5, stress0=-10.0)
    assert out == [(1, 0)]
    _, out = _elastic_run(CASES[0][1], 1000, 1.0, 5, stress0=-0.4)
    assert len(out) == 5 and all(rc == 0 for rc, _ in out)


@pytest.mark.parametrize("name,kw", [CASES[1], CASES[3]], ids=["2d_shear", "3d_shear"])
def test_shear_instability_terminates_on_zero_stress(name, kw):
    # |<sigma_xy>| < EPSILON after the step (STZ/bmg.c:402): start at -2 mu inc, which the
    # elastic increment 2 mu inc brings to exactly zero
    o = _uniform_state(kw, 3.0)
    mu = _moduli(o)["mu"]
    _, out = _elastic_run(kw, 1000, 1.0, 5, stress0=-(2 * mu * 1e-4))
    assert out == [(1, 0)]
    _, out = _elastic_run(kw, 1000, 1.0, 5, stress0=-(2 * mu * 1e-4) * 1.5)
    assert len(out) == 5 and all(rc == 0 for rc, _ in out)


def test_print_restores_initial_barriers():
    """A crystal-like transform redraws the voxel's barriers (STZ/elastic.c:446-506); the
    next print block puts every barrier back to its initial draw (STZ/bmg.c:215-223)."""
    kw = dict(dim=2, n=(16, 16, 1), modes=20, seed=21, loading=0, mro_fraction=1.0, mro_q0=1.5,
              random_ratio=0.2)
    o = po.Oracle(**kw)
    assert np.array_equal(o.field(po.F_Q1), o.field(po.F_Q0))    # Q1 = Q0 (STZ/init.c:189)
    # barriers lowered to 0.8-1.0 eV (the stress-free set, as an upload would set it) so that
    # the first steps are thermal events
    rng = np.random.default_rng(2)
    o.set_barriers(0.9 + 0.05 * rng.standard_normal(o.M * o.N))
    q_init = o.field(po.F_Q0).copy()
    o.set_print(1000, 1.0)          # no print within the first 0.099 of strain
    for _ in range(50):
        rc, info = o.step()
        assert rc == 0 and info.printed == 0
        if info.kind == THERMAL:
            break
    else:
        pytest.fail("no thermal event in 50 steps")
    site = info.site
    q = o.field(po.F_Q0).reshape(20, -1)
    # redrawn as xQ0amp * (1 - r + r U), in [1.2, 1.5]
    assert not np.array_equal(q[:, site], q_init.reshape(20, -1)[:, site])
    assert np.all((q[:, site] >= 1.5 * 0.8) & (q[:, site] <= 1.5))
    assert np.array_equal(o.field(po.F_Q1), q_init)
    o.set_print(0, 1.0)             # the next step prints
    rc, info = o.step()
    assert info.printed == 1
    assert np.array_equal(o.field(po.F_Q0), q_init)


# ---------------------------------------------------------------- ElasticEnergy
@pytest.mark.parametrize("dim,dims", [(2, (8, 4, 1)), (3, (4, 4, 2))])
def test_elastic_energy_closed_form(dim, dims):
    """ElasticEnergy (STZ/elastic.c:533-547) = sum over voxels of sigma:C0^-1:sigma / 2, with
    C0^-1 from a numerical inverse of the Voigt stiffness and the tensor contraction's
    doubled shear terms."""
    kw = dict(dim=dim, n=dims, modes=2)
    if dim == 3:
        kw.update(E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    o = po.Oracle(empty=True, **kw)
    N, C = o.N, o.C
    k = _moduli(o)
    rng = np.random.default_rng(dim + 40)
    sig = rng.standard_normal((C, N))
    o.field(po.F_SIGMA)[:] = sig.ravel()
    inv = np.linalg.inv(_stiffness(k["lam"], k["mu"], dim))
    w = np.array([1, 1, 2.0]) if dim == 2 else np.array([1, 1, 1, 2, 2, 2.0])
    expect = 0.5 * np.sum(w[:, None] * sig * (inv @ sig))
    assert abs(o.elastic_energy() - expect) < 1e-13 * abs(expect)
