"""Pins of the oracle's rate / selection / transform / step logic.

Expected values are computed with numpy from the reference formulas:
ActEnergy STZ/init.c:561-690, SofteningFactor STZ/init.c:711-739,
KMC STZ/evolution.c:102-363, VoxelShearTransform STZ/elastic.c:361-520.
"""
import math

import numpy as np
import pytest

import pyoracle as po

E_J = 1.60217733e-19
KB = 1.380658e-23


def _fill_random(o, rng, sigma_scale=1.0):
    N, C, S, M = o.N, o.C, o.S, o.M
    o.field(po.F_SIGMA)[:] = rng.standard_normal(C * N) * sigma_scale
    o.field(po.F_DEPS)[:] = rng.standard_normal(M * S * N) * 0.07
    o.field(po.F_DEPS_LAST)[:] = rng.standard_normal(S * N) * 0.07
    o.set_barriers(2.0 + 0.25 * rng.standard_normal(M * N))
    o.field(po.F_PERM)[:] = rng.uniform(0, 0.5, N)
    o.field(po.F_TEMP)[:] = rng.uniform(0, 2.0, N)
    o.field(po.F_TLAST)[:] = -rng.uniform(0, 5e-7, N)


def _expand(st, dim):
    if dim == 2:
        return np.stack([st[0], -st[0], st[1]])
    return np.stack([st[0], st[1], -(st[0] + st[1]), st[2], st[3], st[4]])


def _wts(dim):
    return np.array([1, 1, 2.0]) if dim == 2 else np.array([1, 1, 1, 2, 2, 2.0])


@pytest.mark.parametrize("dim,dims", [(2, (8, 4, 1)), (3, (4, 4, 2))])
def test_act_energy_closed_form(dim, dims):
    kw = dict(dim=dim, n=dims, modes=5)
    if dim == 3:
        kw.update(E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    o = po.Oracle(empty=True, **kw)
    rng = np.random.default_rng(dim)
    _fill_random(o, rng)
    totime = 3e-8
    o.set_scalars(np.zeros(6), np.zeros(6), totime)
    o.softening()
    nneg = o.act_energy()
    N, C, S, M = o.N, o.C, o.S, o.M
    T = 300.0
    kT = T * KB * (1 / E_J)
    tTemp = 1 / (1e13 * math.exp(-0.37 / kT))
    soft = o.field(po.F_PERM) + o.field(po.F_TEMP) * np.exp((o.field(po.F_TLAST) - totime) / tTemp)
    assert np.allclose(o.field(po.F_SOFT), soft, rtol=1e-15, atol=0)
    sig = o.field(po.F_SIGMA).reshape(C, N)
    deps = o.field(po.F_DEPS).reshape(M, S, N)
    last = _expand(o.field(po.F_DEPS_LAST).reshape(S, N), dim)
    w = _wts(dim)[:, None]
    V = 0.7 * 0.7 * 1.0
    Q = np.zeros((M, N))
    for m in range(M):
        d = _expand(deps[m], dim)
        W = np.sum(w * sig * d, axis=0) * V * 1e-18 / E_J
        dot = np.sum(w * last * d, axis=0)
        n1 = np.sqrt(np.sum(w * d * d, axis=0))
        n2 = np.sqrt(np.sum(w * last * last, axis=0))
        dm = dot / (n1 + 1e-17) / (n2 + 1e-17)
        D = (1 + dm) ** 2 / 4
        Q[m] = o.field(po.F_Q0).reshape(M, N)[m] * np.exp(-D * soft) - W / 2
    got = o.field(po.F_Q).reshape(M, N)
    assert np.max(np.abs(got - Q)) < 1e-13 * np.max(np.abs(Q))
    assert nneg == int(np.sum(got < 0))


def test_act_energy_initial_state_direction_factor_quarter():
    # with no previous transformation, deps_last = 0 -> D = 1/4 and softening 0 -> Q = Q0 - W/2
    o = po.Oracle(dim=2, n=(8, 8, 1), modes=3, seed=5)
    o.softening()
    o.act_energy()
    q = o.field(po.F_Q)
    assert np.array_equal(q, o.field(po.F_Q0))      # sigma == 0 initially -> W == 0


def test_select_matches_cumsum_search():
    o = po.Oracle(empty=True, dim=2, n=(8, 8, 1), modes=4)
    rng = np.random.default_rng(17)
    q = o.field(po.F_Q)
    q[:] = 1.2 + 0.15 * rng.standard_normal(q.size)
    R = o.total_rate()
    kT = 300.0 * KB * (1 / E_J)
    rate = (0.7 * 0.7 * 1.0 * 1e13) * np.exp(-q / kT)
    assert abs(R - rate.sum()) < 1e-12 * R
    for dice in list(rng.uniform(size=200)) + [1e-12, 0.5, 0.999999]:
        o.total_rate()
        rc, site, mode = o.select(dice)
        S = np.cumsum(rate / rate.sum())
        ev = int(np.searchsorted(S, dice, side="left"))
        assert rc == 0
        assert (mode, site) == (ev // o.N, ev % o.N)


def test_select_failure_codes():
    o = po.Oracle(empty=True, dim=2, n=(4, 4, 1), modes=2)
    o.field(po.F_Q)[:] = 1.0
    o.total_rate()
    assert o.select(0.0)[0] == -1          # dice == 0 is not located (STZ/evolution.c:186-234)
    o.total_rate()
    assert o.select(1.5)[0] == -1          # beyond the normalised total


def test_transform_hand_case():
    o = po.Oracle(empty=True, dim=2, n=(2, 2, 1), modes=2, seed=31)
    N, S, M = o.N, o.S, o.M
    deps = o.field(po.F_DEPS)
    deps[:] = np.arange(M * S * N) * 0.01 + 0.01     # [m][s][v]
    o.field(po.F_PERM)[:] = [0.0, 0.0, 0.5, 0.0]
    o.set_scalars(np.zeros(6), np.zeros(6), 7.5)
    mm, site = 1, 2
    d_xx = deps[(mm * S + 0) * N + site]
    d_xy = deps[(mm * S + 1) * N + site]
    o.transform(mm, site)
    eps0 = o.field(po.F_EPS0).reshape(3, N)
    assert eps0[0, site] == d_xx and eps0[1, site] == -d_xx and eps0[2, site] == d_xy
    assert np.all(eps0[:, [0, 1, 3]] == 0)
    last = o.field(po.F_DEPS_LAST).reshape(S, N)
    assert last[0, site] == d_xx and last[1, site] == d_xy
    ds = d_xx * d_xx + d_xx * d_xx + 2 * d_xy * d_xy
    assert o.field(po.F_PERM)[site] == min(0.5 + ds * 13, -math.log(0.6))
    assert o.field(po.F_TEMP)[site] == ds * 39
    assert o.field(po.F_TLAST)[site] == 7.5
    # all modes of the voxel redrawn from the stream: 0.5*gamma0*G, (xx, xy) per mode
    raw = iter(po.mt_stream(31, 200).tolist())
    from test_oracle_rng import _polar_from_uniforms
    g = [_polar_from_uniforms(raw) for _ in range(2 * M)]
    new = deps.reshape(M, S, N)[:, :, site]
    assert np.array_equal(new.ravel(), np.array([(0.5 * 0.14) * x for x in g]))
    # other voxels untouched
    assert deps.reshape(M, S, N)[0, 0, 0] == 0.01


def _trace(o, nsteps):
    out = []
    for _ in range(nsteps):
        rc, info = o.step()
        assert rc >= 0
        out.append((info.kind, info.site, info.mode, info.n_transformed, info.totime, info.stress[2]))
        if rc == 1:
            break
    return out


def test_trajectory_deterministic_and_sane():
    kw = dict(dim=2, n=(32, 32, 1), modes=20, seed=123, loading=1)
    a = po.Oracle(**kw)
    b = po.Oracle(**kw)
    ta, tb = _trace(a, 400), _trace(b, 400)
    assert ta == tb
    kinds = [t[0] for t in ta]
    # the first steps are purely elastic loading until the stress activates STZs
    assert kinds[0] == 0
    assert 1 in kinds
    times = [t[4] for t in ta]
    assert all(t1 >= t0 for t0, t1 in zip(times, times[1:]))
    for k, site, mode, ntr, _, _ in ta:
        if k == 1:
            assert 0 <= site < 1024 and 0 <= mode < 20
        if k == 2:
            assert ntr >= 1
    c = a.counters()
    assert c[0] == len(ta)


def test_thread_count_does_not_change_results():
    kw = dict(dim=3, n=(16, 16, 8), modes=20, seed=77, loading=1, E=94.5, nu=0.35, double_nu=0,
              gamma0=0.1, q0_std=0.2, athermal_correction=1)
    traces = []
    for th in (1, 4):
        po.set_threads(th)
        o = po.Oracle(**kw)
        o.set_print(5, 1.0)
        traces.append((_trace(o, 300), o.field(po.F_SIGMA).copy()))
    po.set_threads(0)
    assert traces[0][0] == traces[1][0]
    assert np.array_equal(traces[0][1], traces[1][1])
