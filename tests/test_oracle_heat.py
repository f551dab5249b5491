"""CPU pins of the oracle's local temperature (Tlocal) and HeatConducting
(ora_heat_conduct: STZ/elastic.c:549-582; heating in VoxelShearTransform,
STZ/elastic.c:381-383; calls after thermal and elastic steps, STZ/evolution.c:247-249,
325-327).

- A single Fourier mode of the reference grid decays by exactly exp(-|k| dt), |k| from
  the reference's k vectors (k = 2*pi*(j/n - round(j/n)) with its truncated pi), and the
  mean by exp(-dt) (the reference sets |k| = 1 at k = 0).  A |k|^2 kernel, a dropped
  k = 0 quirk or a transposed axis fails it.
- The heating of a transform equals |sigma:deps| * 1e9 / (420 * 6125) computed by hand
  from the voxel's stress and the chosen mode's strain before its redraw.
- With the flag on, a run's Tlocal equals the flag-off run's heating increments passed
  through a numpy FFT restatement of the diffusion after every thermal and elastic step,
  and the event sequence is unchanged (Tlocal feeds nothing back).
"""
import numpy as np
import pytest

import pyoracle as po

PI_REF = 3.14159265359
COEFF = 1 / 420.0 / 6125.0
DIFF = 3e-6


def _kw(dims, **extra):
    if len(dims) == 2:
        kw = dict(dim=2, n=(dims[0], dims[1], 1))
    else:
        kw = dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)
    kw.update(extra)
    return kw


def _kmag(o):
    """|k| on the grid, shape (n2, n1, n0), 1 at k = 0."""
    n0, n1, n2 = o.dims
    ks = []
    for a, n in enumerate((n0, n1, n2)):
        j = np.arange(n)
        s = j * (1.0 / n)
        k1 = (s - np.round(s)) * 2 * PI_REF        # |.| is all that enters; half-way ties agree in size
        sl = [None, None, None]
        sl[2 - a] = slice(None)
        ks.append(np.broadcast_to(k1[tuple(sl)], (n2, n1, n0)) if n > 1 else np.zeros((n2, n1, n0)))
    g2 = np.sqrt(sum(k * k for k in ks[: o.params.dim]))
    g2[0, 0, 0] = 1.0
    return g2


def _diffuse(o, T, dt):
    n0, n1, n2 = o.dims
    f = np.fft.fftn(T.reshape(n2, n1, n0)) * np.exp(-_kmag(o) * dt)
    return np.real(np.fft.ifftn(f)).reshape(-1)


@pytest.mark.parametrize("dims,mode", [((16, 8), (3, 1, 0)), ((16, 8), (8, 4, 0)), ((8, 8, 8), (1, 2, 3)),
                                      ((8, 4, 16), (4, 2, 8))])
def test_single_mode_decays_by_exp_minus_k_dt(dims, mode):
    o = po.Oracle(empty=True, **_kw(dims))
    n0, n1, n2 = o.dims
    i2, i1, i0 = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    ph = 2 * np.pi * (mode[0] * i0 / n0 + mode[1] * i1 / n1 + mode[2] * i2 / max(n2, 1))
    T0, A = 300.0, 5.0
    o.field(po.F_TLOCAL)[:] = (T0 + A * np.cos(ph)).reshape(-1)
    dt = 0.37
    assert o.heat_conduct(dt) == 0
    kk = []
    for j, n in zip(mode, (n0, n1, n2)):
        s = j * (1.0 / n)
        kk.append((s - np.round(s)) * 2 * PI_REF)
    km = np.sqrt(sum(x * x for x in kk[: o.params.dim]))
    want = T0 * np.exp(-dt) + A * np.exp(-km * dt) * np.cos(ph)
    assert np.max(np.abs(o.field(po.F_TLOCAL) - want.reshape(-1))) < 1e-11


def test_random_field_matches_numpy_restatement():
    o = po.Oracle(empty=True, **_kw((16, 8, 4)))
    rng = np.random.default_rng(1)
    T = 300 + rng.standard_normal(o.N)
    o.field(po.F_TLOCAL)[:] = T
    o.heat_conduct(0.05)
    assert np.max(np.abs(o.field(po.F_TLOCAL) - _diffuse(o, T, 0.05))) < 1e-11


@pytest.mark.parametrize("dims", [(16, 8), (8, 4, 8)])
def test_transform_heats_by_sigma_deps(dims):
    o = po.Oracle(empty=True, **_kw(dims))
    rng = np.random.default_rng(2)
    o.field(po.F_SIGMA)[:] = rng.standard_normal(o.C * o.N) * 2.0
    o.field(po.F_DEPS)[:] = rng.standard_normal(o.M * o.S * o.N) * 0.05
    o.field(po.F_TLOCAL)[:] = 300.0
    site, mode = 5, 1
    sg = o.field(po.F_SIGMA).reshape(o.C, o.N)[:, site].copy()
    st = o.field(po.F_DEPS).reshape(o.M, o.S, o.N)[mode, :, site].copy()
    if o.C == 3:
        d, w = np.array([st[0], -st[0], st[1]]), np.array([1, 1, 2.0])
    else:
        d = np.array([st[0], st[1], -(st[0] + st[1]), st[2], st[3], st[4]])
        w = np.array([1, 1, 1, 2, 2, 2.0])
    o.transform(mode, site)
    want = 300.0 + abs(np.sum(w * sg * d) * 1e9) * COEFF
    T = o.field(po.F_TLOCAL)
    assert abs(T[site] - want) < 1e-12 * want
    assert np.all(np.delete(T, site) == 300.0)


@pytest.mark.parametrize("dims,nsteps", [((32, 32), 300), ((8, 8, 8), 600)])
def test_heat_flag_diffuses_after_thermal_and_elastic_steps(dims, nsteps):
    kw = _kw(dims, seed=11, loading=1)
    off = po.Oracle(**kw)
    on = po.Oracle(**kw)
    on.set_heat(True)
    dt_el = kw.get("strainincrement", 1e-4) / kw.get("strainrate", 1e-4)
    T_prev_off = off.field(po.F_TLOCAL).copy()
    T_exp = on.field(po.F_TLOCAL).copy()
    kinds = np.zeros(3, int)
    for i in range(nsteps):
        rc0, a = off.step()
        rc1, b = on.step()
        assert (a.kind, a.site, a.mode, a.n_transformed) == (b.kind, b.site, b.mode, b.n_transformed)
        kinds[a.kind] += 1
        T_off = off.field(po.F_TLOCAL).copy()
        T_exp = T_exp + (T_off - T_prev_off)
        if a.kind == 1:
            T_exp = _diffuse(on, T_exp, DIFF * a.dt_kmc)
        elif a.kind == 0:
            T_exp = _diffuse(on, T_exp, DIFF * dt_el)
        T_prev_off = T_off
        assert np.max(np.abs(on.field(po.F_TLOCAL) - T_exp)) < 1e-9 * np.max(np.abs(T_exp)), i
        if rc0:
            break
    assert kinds[0] > 0 and kinds[1] > 0
