"""CPU pins of the oracle's crystal-like MRO (medium-range order) restatement:
the MRO map (InitMRO 0: STZ/init.c:350-397; InitMRO 1: :398-431), DStrain's families
(STZ/elastic.c:60-138), ActEnergy's NO_SF_XTAL_MRO factors (STZ/init.c:645-667) and
VoxelShearTransform's barrier redraw (STZ/elastic.c:387-396, 446-506).

The random stream is predicted independently of the oracle: raw MT19937 words from
numpy's legacy RandomState (pinned equal to GSL's generator in test_oracle_rng.py),
uniforms w/2^32 and the reference's polar Gaussian on them.  The whole initial state
(map, orientations, barriers, every mode's strain) must then come out bit for bit, which
pins the number and order of draws per voxel and mode.  The formulas for the rotated
family modes are the reference's text (it cannot be built here); the two uniforms inside
each of its expressions are taken left operand first, the order g++ -O3 -std=c++11
(STZ/Makefile:1-3) evaluates a stand-alone expression of the same shape in, and the
rotation's cos/sin pair comes from one sincos call, as g++ compiles that shape (glibc's
sincos can differ from sin() in the last bit).
Closed forms: mode 0 of a family is amp*(cos, sin) of its orientation, modes from the
family size on are zero, the softening factors compound as f^(m+1) per mode.
"""
import ctypes
import math

import numpy as np
import pytest

import pyoracle as po

PI = 3.14159265359
_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_LIBM.sincos.restype = None


def _sincos(x):
    """libm's sincos: g++ computes cos(2*theta), sin(2*theta) of one local with one call."""
    s, c = ctypes.c_double(), ctypes.c_double()
    _LIBM.sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


XM = {1: 6, 2: 2, 3: 12}
FAC = {1: 0.1, 2: 0.01, 3: 0.9}


class Stream:
    """GSL mt19937 from numpy's legacy generator: uniform() = word / 2^32, the polar
    Gaussian of the reference on positive uniforms."""

    def __init__(self, seed=None, state=None):
        self.rs = np.random.RandomState(seed if seed is not None else 0)
        if state is not None:
            self.rs.set_state(("MT19937", np.array(state[0], dtype=np.uint32), int(state[1])))
        self.buf, self.i = [], 0

    def word(self):
        if self.i == len(self.buf):
            self.buf = self.rs.randint(0, 2**32, size=4096, dtype=np.uint32).tolist()
            self.i = 0
        w = self.buf[self.i]
        self.i += 1
        return w

    def uniform(self):
        return self.word() / 4294967296.0

    def upos(self):
        while True:
            x = self.uniform()
            if x != 0:
                return x

    def gaussian(self, sigma):
        while True:
            x = -1 + 2 * self.upos()
            y = -1 + 2 * self.upos()
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0):
                break
        return sigma * y * math.sqrt(-2.0 * math.log(r2) / r2)


def _family(st, t, theta, g0, m, a0, a2):
    """(xx, xy) of mode m of a crystal-like voxel of type t (STZ/elastic.c:62-117)."""
    if m == 0:
        amp = 0.8 * g0 if t == 2 else 0.5 * g0
        return amp * math.cos(theta), amp * math.sin(theta)
    if m < XM[t]:
        th = m * 2. * PI / XM[t] * (1. + 0.2 * (2 * st.uniform() - 1.))
        s2, c2 = _sincos(2 * th)
        u1, u2 = st.uniform(), st.uniform()
        xx = a0 * (1. + 0.2 * (2 * u1 - 1.)) * c2 + a2 * (1. + 0.2 * (2 * u2 - 1.)) * s2
        u3, u4 = st.uniform(), st.uniform()
        xy = -a0 * (1. + 0.2 * (2 * u3 - 1.)) * s2 + a2 * (1. + 0.2 * (2 * u4 - 1.)) * c2
        return xx, xy
    return 0.0, 0.0


def _predict_fields(st, types, thetas, kw, M, N):
    q0 = np.zeros((M, N))
    for m in range(M):
        for v in range(N):
            q0[m, v] = st.gaussian(kw["q0_std"]) + kw["q0_mean"]
    deps = np.zeros((M, 2, N))
    g0 = kw["gamma0"]
    for m in range(M):
        for v in range(N):
            t = types[v]
            if t == 0:
                deps[m, 0, v] = 0.5 * g0 * st.gaussian(1.0)
                deps[m, 1, v] = 0.5 * g0 * st.gaussian(1.0)
            else:
                deps[m, 0, v], deps[m, 1, v] = _family(st, t, thetas[v], g0, m, deps[0, 0, v], deps[0, 1, v])
    return q0, deps


def _kw(**extra):
    kw = dict(po.REFERENCE_2D)
    kw.update(dim=2, n=(16, 12, 1), modes=14, seed=7)
    kw.update(extra)
    return kw


def test_random_map_and_initial_state_from_the_stream():
    kw = _kw(mro_fraction=0.3)
    o = po.Oracle(**kw)
    N, M = o.N, o.M
    st = Stream(seed=kw["seed"])
    types, thetas = np.zeros(N, int), np.zeros(N)
    for v in range(N):
        if st.uniform() <= kw["mro_fraction"]:
            types[v] = 1
            thetas[v] = st.uniform() * PI - PI / 2.0
    q0, deps = _predict_fields(st, types, thetas, kw, M, N)
    assert 0.15 * N < types.sum() < 0.45 * N
    assert np.array_equal(o.mro_map(), types)
    assert np.array_equal(o.field(po.F_Q0).reshape(M, N), q0)
    assert np.array_equal(o.field(po.F_DEPS).reshape(M, 2, N), deps)
    # the stream position after the initial state: the next draw of both agrees
    r = o.rng().contents
    nxt = Stream(state=(list(r.mt), r.mti)).word()
    assert nxt == st.word()


def test_file_map_families_and_ten_fold_split():
    kw = _kw()
    rng = np.random.default_rng(3)
    N = kw["n"][0] * kw["n"][1]
    types = rng.integers(0, 4, N)
    thetas = rng.uniform(-PI / 2, PI / 2, N)
    o = po.Oracle(mro_map=(types, thetas), **kw)
    M = o.M
    st = Stream(seed=kw["seed"])
    t2 = types.copy()
    for v in range(N):
        if t2[v] == 3 and st.uniform() < 0.5:
            t2[v] = 2
    assert np.array_equal(o.mro_map(), t2)
    assert (t2 == 3).any() and (t2 == 2).any()
    q0, deps = _predict_fields(st, t2, thetas, kw, M, N)
    assert np.array_equal(o.field(po.F_Q0).reshape(M, N), q0)
    got = o.field(po.F_DEPS).reshape(M, 2, N)
    assert np.array_equal(got, deps)
    # closed forms: mode 0 along the orientation with the family's amplitude, zero
    # strain from the family size on
    g0 = kw["gamma0"]
    for v in range(N):
        t = t2[v]
        if t == 0:
            continue
        amp = (0.8 if t == 2 else 0.5) * g0
        assert got[0, 0, v] == amp * math.cos(thetas[v]) and got[0, 1, v] == amp * math.sin(thetas[v])
        assert np.all(got[XM[t]:, :, v] == 0.0)
        for m in range(1, min(XM[t], M)):
            ratio = math.hypot(got[m, 0, v], got[m, 1, v]) / amp
            assert 0.5 < ratio < 1.7


def test_act_energy_factors_compound_over_modes():
    kw = _kw(n=(8, 8, 1), modes=6, softening=1)
    rng = np.random.default_rng(5)
    N = 64
    types = rng.integers(0, 4, N)
    o = po.Oracle(mro_map=(types, rng.uniform(-1, 1, N)), **kw)
    t = o.mro_map()
    o.field(po.F_SIGMA)[:] = rng.standard_normal(3 * N)
    o.field(po.F_DEPS_LAST)[:] = rng.standard_normal(2 * N) * 0.05
    o.field(po.F_PERM)[:] = rng.uniform(0.1, 0.5, N)
    o.softening()
    S0 = o.field(po.F_SOFT).copy()
    o.act_energy()
    q = o.field(po.F_Q).reshape(6, N)
    q0 = o.field(po.F_Q0).reshape(6, N)
    deps = o.field(po.F_DEPS).reshape(6, 2, N)
    sg = o.field(po.F_SIGMA).reshape(3, N)
    pl = o.field(po.F_DEPS_LAST).reshape(2, N)
    voxel = kw["lx"] * kw["ly"] * kw["lz"]
    w = np.array([1, 1, 2.0])
    for m in range(6):
        d = np.stack([deps[m, 0], -deps[m, 0], deps[m, 1]])
        p = np.stack([pl[0], -pl[0], pl[1]])
        dwork = (w[:, None] * sg * d).sum(0) * voxel * 1e-18 / 1.60217733e-19
        dm = (w[:, None] * p * d).sum(0) / (np.sqrt((w[:, None] * d * d).sum(0)) + 1e-17) / \
            (np.sqrt((w[:, None] * p * p).sum(0)) + 1e-17)
        dm = (1 + dm) ** 2 / 4
        f = np.array([FAC.get(int(x), 1.0) for x in t])
        soft = np.where(t > 0, S0 * f ** (m + 1), S0)
        want = q0[m] * np.exp(-(dm * np.where(t > 0, f, 1.0)) * soft) - dwork / 2
        assert np.allclose(q[m], want, rtol=1e-13, atol=1e-13), m


@pytest.mark.parametrize("t", [0, 1, 2, 3])
def test_transform_of_each_type(t):
    kw = _kw(n=(8, 8, 1), modes=5, mro_q0=2.5, random_ratio=0.3)
    N = 64
    types = np.zeros(N, int)
    types[9] = t
    o = po.Oracle(mro_map=(types, np.full(N, 0.4)), **kw)
    o.field(po.F_SIGMA)[:] = 1.0
    deps0 = o.field(po.F_DEPS).copy()
    q00 = o.field(po.F_Q0).copy()
    perm0, v, mode = o.field(po.F_PERM).copy(), 9, 1
    r = o.rng().contents
    st = Stream(state=(list(r.mt), r.mti))
    d = deps0.reshape(5, 2, N)[mode, :, v]
    ds = d[0] * d[0] + (-d[0]) * (-d[0]) + (2 * d[1]) * d[1]
    o.transform(mode, v)
    six = (o.mro_map()[v] == 1)
    assert o.field(po.F_PERM)[v] == perm0[v] + ds * (20.0 if six else 13.0) or o.field(po.F_PERM)[v] == kw["perm_cap"]
    assert o.field(po.F_TEMP)[v] == ds * (60.0 if six else 39.0)
    q0 = o.field(po.F_Q0).reshape(5, N)
    deps = o.field(po.F_DEPS).reshape(5, 2, N)
    if o.mro_map()[v] == 0:
        assert np.array_equal(q0, q00.reshape(5, N))
        want = np.array([[0.5 * kw["gamma0"] * st.gaussian(1.0) for _ in range(2)] for _ in range(5)])
        assert np.array_equal(deps[:, :, v], want)
    else:
        assert np.array_equal(deps, deps0.reshape(5, 2, N))
        want = np.array([2.5 * (1 - 0.3 + 0.3 * (st.uniform())) for _ in range(5)])
        assert np.array_equal(q0[:, v], want)
        assert np.all((want >= 2.5 * 0.7) & (want <= 2.5))


def test_mro_trajectory_transforms_crystal_voxels():
    kw = _kw(n=(32, 32, 1), modes=20, mro_fraction=0.3, seed=21, loading=1, mro_q0=1.5)
    o = po.Oracle(**kw)
    o.set_print(5, 1.0)
    hits = 0
    for i in range(1500):
        rc, info = o.step()
        if info.kind == 1 and info.mro:
            hits += 1
            q0 = o.field(po.F_Q0).reshape(20, -1)[:, info.site]
            assert np.all((q0 >= 1.5 * 0.8) & (q0 <= 1.5))
        if rc:
            break
    assert hits > 0
