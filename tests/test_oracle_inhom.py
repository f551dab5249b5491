"""CPU pins of the oracle's elastically inhomogeneous stress solve (ora_set_moduli +
ora_stress_solve: the restatement of InhomogeneousStressSolver, STZ/elastic.c:977-1066,
with non-uniform Lambda(x), Mu(x) in place of STZ/init.c:457-465's uniform ones).

- Uniform moduli given explicitly take the fixed-point loop and must reproduce the
  uniform path bit for bit (one homogeneous solve: the reference's exit after the first
  iteration when Lambda = lambda, Mu = mu).
- A laminate (moduli varying along one grid axis only) has an exact solution that the
  fixed point must reach to its stopping tolerance: the tangential strains equal the
  mean strain, the tractions on the layer planes are uniform, and the normal strains
  follow from layer-wise elasticity (textbook laminate homogenisation, solved here with
  full 3x3x3x3 isotropic stiffness tensors in numpy).  Shear along the layers gives the
  harmonic mean of mu, shear across them the arithmetic mean, so an axis mix-up, a sign
  error in (C0 - C(x)) or a dropped C(x):eps0 term fails it.  The loop stops once the
  RMS change of its eigenstrain is below 1e-3 |mean strain|, which leaves errors of up
  to about 5e-3 of the largest stress: the tolerance is 1e-2 of it, while the means
  above differ by 16% and the eigenstrain moves the stresses by 40%.
- A strong contrast that the 0.5-mixed iteration cannot settle stops at the reference's
  cap: 101 homogeneous solves (break on step > 100).
Tensor index a of a strain component pairs with storage axis a (axis 0 the fastest),
the pairing the displacement tests pin through the spectral gradient.
"""
import numpy as np
import pytest

import pyoracle as po

COMP = {2: [(0, 0), (1, 1), (0, 1)], 3: [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]}


def _kw(dims):
    if len(dims) == 2:
        return dict(dim=2, n=(dims[0], dims[1], 1))
    return dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)


def _layers(o, axis, values):
    """Per-voxel array that takes values[i % len(values)] along storage axis `axis`."""
    n0, n1, n2 = o.dims
    i2, i1, i0 = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    idx = (i0, i1, i2)[axis]
    return np.asarray(values, dtype=np.float64)[idx % len(values)].reshape(-1)


def _stiffness(lam, mu, d):
    I = np.eye(d)
    return (lam * np.einsum("ij,kl->ijkl", I, I)
            + mu * (np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I)))


def _to_tensor(v, d):
    t = np.zeros((d, d))
    for c, (a, b) in enumerate(COMP[d]):
        t[a, b] = t[b, a] = v[c]
    return t


def _laminate(lams, mus, axis, d, ebar, e0):
    """Exact strain and stress of each distinct layer of an equal-thickness laminate with
    normal `axis`, mean total strain `ebar` and uniform eigenstrain `e0` (plane strain in
    2-D).  Returns the stresses sigma(layer) as d x d tensors."""
    Eb, E0 = _to_tensor(ebar, d), _to_tensor(e0, d)
    n = axis
    Cs = [_stiffness(l, m, d) for l, m in zip(lams, mus)]
    # unknowns per layer: the normal column of the strain, eps[n, :] (= eps[:, n]).
    # eps = Eb + sym(dv (x) e_n) with <dv> = 0; traction t = sigma[n, :] uniform.
    # sigma[n, :] = C[n, :, k, l] (eps - E0)[k, l] is affine in dv: A_L dv + b_L = t.
    A, b = [], []
    en = np.eye(d)[n]
    for C in Cs:
        Ai = np.zeros((d, d))
        for q in range(d):
            dv = np.eye(d)[q]
            de = 0.5 * (np.outer(dv, en) + np.outer(en, dv))
            Ai[:, q] = np.einsum("jkl,kl->j", C[n], de)
        A.append(Ai)
        b.append(np.einsum("jkl,kl->j", C[n], Eb - E0))
    Ainv = [np.linalg.inv(a) for a in A]
    # dv_L = Ainv_L (t - b_L), mean over layers = 0  ->  t = <Ainv>^-1 <Ainv b>
    t = np.linalg.solve(np.mean(Ainv, axis=0), np.mean([ai @ bi for ai, bi in zip(Ainv, b)], axis=0))
    out = []
    for C, ai, bi in zip(Cs, Ainv, b):
        dv = ai @ (t - bi)
        eps = Eb + 0.5 * (np.outer(dv, en) + np.outer(en, dv))
        out.append(np.einsum("ijkl,kl->ij", C, eps - E0))
    return out


def _solve_uniform_and_explicit(dims, seed):
    res = []
    for explicit in (False, True):
        o = po.Oracle(empty=True, **_kw(dims))
        rng = np.random.default_rng(seed)
        o.field(po.F_EPS0)[:] = rng.standard_normal(o.C * o.N) * 0.03
        ea = rng.standard_normal(6) * 0.01
        o.set_scalars(ea, np.zeros(6), 0.0)
        if explicit:
            o.set_moduli(np.full(o.N, o.consts[3]), np.full(o.N, o.consts[2]))
        o.solve()
        res.append((o.field(po.F_SIGMA).copy(), o.scalars()[1].copy(), o.inhom_iterations()))
    return res


@pytest.mark.parametrize("dims", [(16, 8), (8, 4, 16)])
def test_explicit_uniform_moduli_take_one_solve_bit_for_bit(dims):
    (s0, m0, it0), (s1, m1, it1) = _solve_uniform_and_explicit(dims, 5)
    assert it0 == 1 and it1 == 1
    assert np.array_equal(s0, s1)
    assert np.array_equal(m0, m1)


CASES = [
    # dims, layer axis, strain component loaded (mean strain and eigenstrain)
    ((32, 16), 0, 2), ((32, 16), 1, 2), ((16, 32), 0, 0), ((16, 32), 1, 1),
    ((16, 8, 8), 0, 5), ((8, 8, 16), 2, 5), ((8, 16, 8), 1, 3), ((8, 8, 16), 2, 2),
]


@pytest.mark.parametrize("dims,axis,comp", CASES)
def test_laminate_reaches_the_exact_solution(dims, axis, comp):
    o = po.Oracle(empty=True, **_kw(dims))
    d = o.params.dim
    lam0, mu0 = o.consts[3], o.consts[2]
    fac = np.array([0.6, 1.4, 1.0, 0.8])
    lams, mus = lam0 * fac, mu0 * fac
    o.set_moduli(_layers(o, axis, lams), _layers(o, axis, mus))
    ebar = np.zeros(o.C)
    ebar[comp] = 0.01
    e0 = np.zeros(o.C)
    e0[comp] = 0.004
    e0[(comp + 1) % o.C] = -0.002
    o.field(po.F_EPS0)[:] = np.repeat(e0, o.N)
    ea = np.zeros(6)
    ea[: o.C] = ebar
    o.set_scalars(ea, np.zeros(6), 0.0)
    o.solve()
    it = o.inhom_iterations()
    assert 1 <= it <= 100
    exact = _laminate(lams, mus, axis, d, ebar, e0)
    sig = o.field(po.F_SIGMA).reshape(o.C, o.N)
    lay = _layers(o, axis, np.arange(4)).astype(int)
    scale = max(abs(x) for s in exact for x in s.ravel())
    for L in range(4):
        want = np.array([exact[L][a, b] for a, b in COMP[d]])
        got = sig[:, lay == L]
        err = np.abs(got - want[:, None]).max()
        assert err < 1e-2 * scale, (L, err, scale)
    # the mean stress is the mean of the layers' stresses
    mean = np.array([np.mean([exact[L][a, b] for L in range(4)]) for a, b in COMP[d]])
    assert np.abs(o.scalars()[1][: o.C] - mean).max() < 1e-2 * scale


def test_shear_means_differ_along_and_across_layers():
    """The laminate pin distinguishes harmonic from arithmetic means by far more than
    its tolerance."""
    mu = np.array([0.6, 1.4, 1.0, 0.8])
    harm = 1.0 / np.mean(1.0 / mu)
    assert (np.mean(mu) - harm) / np.mean(mu) > 0.05
    s_across = _laminate(mu * 2, mu, 0, 3, np.array([0, 0, 0, 0, 0, 0.01]), np.zeros(6))
    s_along = _laminate(mu * 2, mu, 2, 3, np.array([0, 0, 0, 0, 0, 0.01]), np.zeros(6))
    assert np.allclose([s[0, 1] for s in s_across], 2 * harm * 0.01)
    assert np.allclose([s[0, 1] for s in s_along], 2 * mu * 0.01)


def test_strong_contrast_stops_at_the_iteration_cap():
    o = po.Oracle(empty=True, **_kw((16, 16)))
    lam0, mu0 = o.consts[3], o.consts[2]
    fac = np.array([1.0, 9.0])
    o.set_moduli(_layers(o, 0, lam0 * fac), _layers(o, 0, mu0 * fac))
    o.set_scalars(np.array([0.0, 0.0, 0.01, 0, 0, 0]), np.zeros(6), 0.0)
    o.solve()
    assert o.inhom_iterations() == 101


def test_small_mean_strain_exits_after_one_solve():
    o = po.Oracle(empty=True, **_kw((16, 16)))
    lam0, mu0 = o.consts[3], o.consts[2]
    fac = np.array([0.6, 1.4])
    o.set_moduli(_layers(o, 0, lam0 * fac), _layers(o, 0, mu0 * fac))
    o.field(po.F_EPS0)[:] = 0.01
    o.set_scalars(np.array([5e-5, 0.0, 0.0, 0, 0, 0]), np.zeros(6), 0.0)
    o.solve()
    assert o.inhom_iterations() == 1
