"""CPU pins of the event-local stress update, on the oracle's full FFT solve.

The GPU's stz_stress_update_local replaces a full solve after transforms at a few sites
by: the field as last solved + C0:(change of EpsAvg) + the superposed impulse responses
of the sites, each the response at voxel 0 shifted periodically to the site.  These tests
check what that rests on, with the oracle's own solve (a restatement of
STZ/elastic.c:779-828 behind the wrapper STZ/elastic.c:977-1066):

- translation invariance: the response to an impulse at site s is the response at
  voxel 0 rolled by s;
- linearity: a multi-site eigenstrain change plus a mean-strain change re-solved in full
  equals the old field + C0:(mean-strain change) + the shifted responses, written here
  in numpy independently of the HIP kernel;
- which symmetry halves the response table: the response is even, K(-x) = K(x) (the
  reference's Re() averages the multiplier at k and -k), but it is not symmetric under
  the reflection of one axis: the reference's round() k vectors (STZ/init.c:126-129) put
  every Nyquist component at -pi, which leaves a (-1)^(x+y) checkerboard on wavevectors
  with two Nyquist components, so an octant-folded table would be wrong.
"""
import numpy as np
import pytest

import pyoracle as po

COMP = {2: ["xx", "yy", "xy"], 3: ["xx", "yy", "zz", "yz", "xz", "xy"]}
# stored transformation-strain components and the tensors they stand for
BASIS = {2: [{"xx": 1.0, "yy": -1.0}, {"xy": 1.0}],
         3: [{"xx": 1.0, "zz": -1.0}, {"yy": 1.0, "zz": -1.0}, {"yz": 1.0}, {"xz": 1.0}, {"xy": 1.0}]}


def _kw(dims):
    if len(dims) == 2:
        return dict(dim=2, n=(dims[0], dims[1], 1))
    return dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)


def _shape(o):
    n0, n1, n2 = o.dims
    return (o.C, n2, n1, n0)


def _solve(o, eps0, epsavg):
    o.field(po.F_EPS0)[:] = eps0.ravel()
    o.set_scalars(epsavg, np.zeros(6), 0.0)
    o.solve()
    return o.field(po.F_SIGMA).reshape(_shape(o)).copy()


def _basis_field(o, k, site):
    dim = o.params.dim
    e = np.zeros(_shape(o))
    n0, n1, _ = o.dims
    i0, i1, i2 = site % n0, (site // n0) % n1, site // (n0 * n1)
    for lab, val in BASIS[dim][k].items():
        e[COMP[dim].index(lab), i2, i1, i0] = val
    return e


def _responses(o):
    return [_solve(o, _basis_field(o, k, 0), np.zeros(6)) for k in range(o.S)]


def _c0(o, de):
    lam, mu = o.consts[3], o.consts[2]
    if o.C == 3:
        tr = de[0] + de[1]
        return np.array([lam * tr + 2 * mu * de[0], lam * tr + 2 * mu * de[1], 2 * mu * de[2]])
    tr = de[0] + de[1] + de[2]
    return np.array([lam * tr + 2 * mu * de[c] if c < 3 else 2 * mu * de[c] for c in range(6)])


GRIDS = [(16, 8), (8, 8, 8), (8, 4, 16)]


@pytest.mark.parametrize("dims", GRIDS)
def test_impulse_response_is_translation_invariant(dims):
    o = po.Oracle(empty=True, **_kw(dims))
    K = _responses(o)
    n0, n1, n2 = o.dims
    rng = np.random.default_rng(7)
    for _ in range(3):
        site = int(rng.integers(o.N))
        sh = (site // (n0 * n1), (site // n0) % n1, site % n0)
        for k in range(o.S):
            r = _solve(o, _basis_field(o, k, site), np.zeros(6))
            rolled = np.roll(K[k], shift=sh, axis=(1, 2, 3))
            assert np.max(np.abs(r - rolled)) <= 1e-13 * np.max(np.abs(K[k]))


@pytest.mark.parametrize("dims", GRIDS)
def test_impulse_response_is_even(dims):
    """K(-x) = K(x) in every component: the reference keeps Re() of a full complex
    inverse transform (STZ/elastic.c:811-824), which applies the average of the
    multiplier at k and -k, an even operator -- so half of the response table (offsets
    with the slowest axis in [0, n/2]) determines the rest."""
    o = po.Oracle(empty=True, **_kw(dims))
    for K in _responses(o):
        inv = np.roll(np.flip(K, axis=(1, 2, 3)), 1, axis=(1, 2, 3))   # x -> (-x) mod n on every axis
        assert np.max(np.abs(inv - K)) <= 1e-13 * np.max(np.abs(K))


def test_reflection_asymmetry_is_the_double_nyquist_checkerboard():
    """2-D: reflecting x maps the xx response of the xy eigenstrain to minus itself except
    for a (-1)^(x+y) term -- the k = (N0/2, N1/2) bin, where n = (-pi, -pi)/|.| has
    n_x n_y > 0 while a reflection-symmetric grid would average it to zero."""
    o = po.Oracle(empty=True, **_kw((16, 8)))
    K = _responses(o)
    kxx = K[1][0, 0]                                   # (y, x)
    refl = np.roll(np.flip(kxx, axis=1), 1, axis=1)    # x -> (-x) mod n
    diff = refl + kxx                                  # zero for an odd function of x
    y, x = np.meshgrid(np.arange(8), np.arange(16), indexing="ij")
    cb = (-1.0) ** (x + y)
    q = float(np.mean(diff * cb))
    assert abs(q) > 1e-3 * np.max(np.abs(kxx))
    assert np.max(np.abs(diff - q * cb)) <= 1e-12 * np.max(np.abs(kxx))


def _shifted_sum(o, K, sites, deps):
    """sum over sites s and basis k of deps[s][k] * K_k((x - s) mod n) (numpy)."""
    n0, n1, _ = o.dims
    out = np.zeros(_shape(o))
    for s, d in zip(sites, deps):
        sh = (s // (n0 * n1), (s // n0) % n1, s % n0)
        for k in range(o.S):
            out += d[k] * np.roll(K[k], shift=sh, axis=(1, 2, 3))
    return out


@pytest.mark.parametrize("dims", GRIDS)
def test_local_update_equals_full_solve(dims):
    o = po.Oracle(empty=True, **_kw(dims))
    K = _responses(o)
    rng = np.random.default_rng(11)
    eps0 = rng.standard_normal(_shape(o)) * 0.05
    if o.C == 6:
        eps0[2] = -(eps0[0] + eps0[1])              # traceless like every transform
    else:
        eps0[1] = -eps0[0]
    ea0 = rng.standard_normal(6) * 0.01
    sig0 = _solve(o, eps0, ea0)
    sites = [0, o.N - 1] + [int(x) for x in rng.integers(o.N, size=3)]
    deps = rng.standard_normal((len(sites), o.S)) * 0.07
    eps1 = eps0.copy()
    for s, d in zip(sites, deps):
        for k in range(o.S):
            eps1 += d[k] * _basis_field(o, k, s)
    ea1 = ea0 + rng.standard_normal(6) * 0.001
    full = _solve(o, eps1, ea1)
    uni = _c0(o, (ea1 - ea0)[: o.C])
    loc = sig0 + uni[:, None, None, None] + _shifted_sum(o, K, sites, deps)
    assert np.max(np.abs(loc - full)) <= 1e-12 * np.max(np.abs(full))


def test_local_update_detects_transposed_axes():
    """The shifted sum with the site's coordinates applied to the wrong axes (a plausible
    indexing slip) misses the full solve."""
    o = po.Oracle(empty=True, **_kw((8, 4, 16)))
    K = _responses(o)
    site = 3 + 8 * 1 + 32 * 11
    d = np.array([[0.0, 0.0, 0.0, 0.0, 0.07]])        # xy eigenstrain
    exact = _solve(o, 0.07 * _basis_field(o, 4, site), np.zeros(6))
    good = _shifted_sum(o, K, [site], d)
    assert np.max(np.abs(good - exact)) <= 1e-12 * np.max(np.abs(exact))
    bad = 0.07 * np.roll(K[4], shift=(3, 1, 11), axis=(1, 2, 3))   # (x, y, z) on (z, y, x)
    assert np.max(np.abs(bad - exact)) > 1e-3 * np.max(np.abs(exact))


@pytest.mark.parametrize("dims", GRIDS + [(32, 16), (16, 16, 16)])
def test_response_table_compaction(dims):
    """The compacted table of the local update (20 planes instead of 30 in 3-D, 5 instead
    of 6 in 2-D) rests on A = T^T M K being symmetric at every offset (T: the stored
    strain components' basis tensors, M: weight 2 on shear components) -- the elastic
    operator's major symmetry -- and on K being recovered from A and the responses of the
    reference normal stress (zz in 3-D, yy in 2-D): sigma_Z = z, the other normal
    components u_i + z, the shear components u_i / 2 with u = A d.  A planted slip (the
    wrong reference row) is caught."""
    o = po.Oracle(**_kw(dims))
    C, S, dim = o.C, o.S, o.params.dim
    K = np.stack([r.reshape(C, -1) for r in _responses(o)], axis=1)      # [C][S][N]
    T = np.zeros((C, S))
    for k in range(S):
        for lab, val in BASIS[dim][k].items():
            T[COMP[dim].index(lab), k] = val
    M = np.diag([1.0, 1.0, 2.0] if C == 3 else [1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    A = np.einsum("ck,cd,dsn->ksn", T, M, K)
    scale = np.abs(K).max()
    assert np.abs(A - A.transpose(1, 0, 2)).max() <= 1e-15 * scale
    A = 0.5 * (A + A.transpose(1, 0, 2))

    zc = 1 if dim == 2 else 2                                   # the reference normal row
    rows = [0, 2] if dim == 2 else [0, 1, 3, 4, 5]              # stress component of each row of A
    nn = 1 if dim == 2 else 2                                   # normal rows of A

    def rebuild(zsrc, half=0.5):
        z = K[zsrc]                                             # [S][N]
        Kr = np.zeros_like(K)
        for i, c in enumerate(rows):
            Kr[c] = A[i] + z if i < nn else half * A[i]
        Kr[zc] = z
        return Kr

    assert np.abs(rebuild(zc) - K).max() <= 1e-14 * scale
    # planted slips: the reference row taken from xx, the shear weight not undone
    assert np.abs(rebuild(0) - K).max() > 1e-3 * scale
    assert np.abs(rebuild(zc, half=1.0) - K).max() > 1e-3 * scale
