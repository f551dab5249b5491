"""Pins of the oracle's Green's-function stress solve (STZ/elastic.c:779-1066)."""
import numpy as np
import pytest

import pyoracle as po
import solve_ref as sr


def _oracle(dim, dims, **kw):
    p = dict(dim=dim, n=dims, modes=2)
    if dim == 3:
        p.update(E=94.5, nu=0.35, double_nu=0)
    p.update(kw)
    return po.Oracle(empty=True, **p)


def _grid_shape(o):
    d = o.dims
    return (d[1], d[0]) if o.params.dim == 2 else (d[2], d[1], d[0])


def _consts(o):
    c = o.consts
    return c[3], c[2]   # lambda, mu


@pytest.mark.parametrize("dim,dims", [(2, (8, 8, 1)), (2, (16, 8, 1)), (3, (8, 8, 8)), (3, (4, 8, 6))])
def test_uniform_eigenstrain_gives_uniform_stress(dim, dims):
    o = _oracle(dim, dims)
    C = o.C
    e = np.linspace(0.01, 0.03, C)
    ea = np.linspace(-0.02, 0.05, 6)
    eps0 = o.field(po.F_EPS0)
    for c in range(C):
        eps0[c * o.N:(c + 1) * o.N] = e[c]
    o.set_scalars(ea, np.zeros(6), 0.0)
    o.solve()
    lam, mu = _consts(o)
    d = sr.to_tensor((ea[:C] - e)[:, None], dim)[0]
    sig = sr.from_tensor(sr.c0_mul(d, lam, mu, dim)[None], dim)[:, 0]
    got = o.field(po.F_SIGMA).reshape(C, o.N)
    assert np.max(np.abs(got - sig[:, None])) < 1e-13 * np.max(np.abs(sig))
    _, stress, _ = o.scalars()
    assert np.allclose(stress[:C], sig, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("dim,dims,kidx", [(2, (16, 16, 1), (3, 2)), (2, (16, 8, 1), (1, 3)),
                                           (3, (8, 8, 8), (1, 2, 3)), (3, (8, 16, 8), (2, 0, 1))])
def test_compatible_eigenstrain_is_stress_free(dim, dims, kidx):
    # eps0 = sym(a (x) n) cos(k.x) is the strain of a displacement field: no fluctuation stress
    o = _oracle(dim, dims)
    C = o.C
    shape = _grid_shape(o)
    ks = sr.kvec(dims, dim)
    kv = np.array([ks[a][kidx[a]] for a in range(dim)])
    n = kv / np.linalg.norm(kv)
    a = np.array([0.3, -0.7, 0.5][:dim])
    t = 0.5 * (np.outer(a, n) + np.outer(n, a))
    coords = np.meshgrid(*[np.arange(dims[a]) for a in reversed(range(dim))], indexing="ij")
    phase = sum(2 * np.pi * kidx[a] * coords[dim - 1 - a] / dims[a] for a in range(dim))
    f = 0.01 * np.cos(phase)
    eps0 = o.field(po.F_EPS0)
    for c, (i, j) in enumerate(sr.pairs(dim)):
        eps0[c * o.N:(c + 1) * o.N] = (t[i, j] * f).ravel()
    o.set_scalars(np.zeros(6), np.zeros(6), 0.0)
    o.solve()
    got = o.field(po.F_SIGMA).reshape(C, o.N)
    lam, mu = _consts(o)
    assert np.max(np.abs(got)) < 1e-13 * 0.01 * (lam + 2 * mu)


@pytest.mark.parametrize("dim,dims", [(2, (16, 16, 1)), (3, (16, 4, 4))])
def test_laminate_closed_form(dim, dims):
    # eps0 varying only along axis 0: traction continuity gives per-voxel closed form
    o = _oracle(dim, dims)
    C = o.C
    lam, mu = _consts(o)
    rng = np.random.default_rng(11)
    prof = rng.standard_normal((C, dims[0])) * 0.01
    shape = _grid_shape(o)
    eps0 = o.field(po.F_EPS0)
    full = np.zeros((C,) + shape)
    for c in range(C):
        full[c] = np.broadcast_to(prof[c], shape)
        eps0[c * o.N:(c + 1) * o.N] = full[c].ravel()
    o.set_scalars(np.zeros(6), np.zeros(6), 0.0)
    o.solve()
    got = o.field(po.F_SIGMA).reshape((C,) + shape)
    # closed form: fluctuating eigenstress s0 = C0:eps0 - <C0:eps0>; the total strain
    # fluctuation is e_xx = s0_xx/(lam+2mu), e_xy = s0_xy/(2mu), e_xz = s0_xz/(2mu)
    s0 = sr.from_tensor(sr.c0_mul(sr.to_tensor(full, dim), lam, mu, dim), dim)
    s0 = s0 - s0.mean(axis=tuple(range(1, dim + 1)), keepdims=True)
    P = sr.pairs(dim)
    eps = np.zeros_like(full)
    for c, (i, j) in enumerate(P):
        if (i, j) == (0, 0):
            eps[c] = s0[c] / (lam + 2 * mu)
        elif i == 0 or j == 0:
            eps[c] = s0[c] / (2 * mu)
    exp = sr.from_tensor(sr.c0_mul(sr.to_tensor(eps - full, dim), lam, mu, dim), dim)
    assert np.max(np.abs(got - exp)) < 1e-12 * np.max(np.abs(exp))
    # the fluctuation of the traction on the laminate normal vanishes
    for c, (i, j) in enumerate(P):
        if i == 0 or j == 0:
            assert np.max(np.abs(got[c] - got[c].mean())) < 1e-12 * np.max(np.abs(exp))


@pytest.mark.parametrize("dim,dims", [(2, (4, 4, 1)), (2, (8, 16, 1)), (3, (4, 4, 4)), (3, (8, 4, 6))])
def test_random_eigenstrain_matches_dense_green_solve(dim, dims):
    o = _oracle(dim, dims)
    C = o.C
    lam, mu = _consts(o)
    shape = _grid_shape(o)
    rng = np.random.default_rng(dims[0] * 7 + dims[1])
    full = rng.standard_normal((C,) + shape) * 0.05
    eps0 = o.field(po.F_EPS0)
    eps0[:] = full.reshape(-1)
    ea = rng.standard_normal(6) * 0.01
    o.set_scalars(ea, np.zeros(6), 0.0)
    o.solve()
    _, sig = sr.green_solve(full, ea, dims, dim, lam, mu)
    got = o.field(po.F_SIGMA).reshape((C,) + shape)
    assert np.max(np.abs(got - sig)) < 1e-12 * np.max(np.abs(sig))
    _, stress, _ = o.scalars()
    assert np.allclose(stress[:C], sig.reshape(C, -1).mean(axis=1), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dim,dims", [(2, (16, 16, 1)), (3, (8, 8, 8))])
def test_equilibrium_k_dot_sigma_hat(dim, dims):
    # the invariant of CheckStressSelfBalance (STZ/elastic.c:625-641), off the Nyquist planes
    o = _oracle(dim, dims)
    C = o.C
    shape = _grid_shape(o)
    rng = np.random.default_rng(2)
    o.field(po.F_EPS0)[:] = (rng.standard_normal((C,) + shape) * 0.05).ravel()
    o.set_scalars(np.zeros(6), np.zeros(6), 0.0)
    o.solve()
    sig = o.field(po.F_SIGMA).reshape((C,) + shape)
    sh = np.fft.fftn(sig, axes=tuple(range(1, dim + 1)))
    T = sr.to_tensor(sh, dim)
    ks = sr.kvec(dims, dim)
    grids = np.meshgrid(*[ks[a] for a in reversed(range(dim))], indexing="ij")
    kv = np.stack([grids[dim - 1 - a] for a in range(dim)], axis=-1)
    res = np.einsum("...ij,...j->...i", T, kv)
    mask = np.ones(shape, bool)
    for a in range(dim):
        sl = [slice(None)] * dim
        sl[dim - 1 - a] = dims[a] // 2
        mask[tuple(sl)] = False
    assert np.max(np.abs(res[mask])) < 1e-10 * np.max(np.abs(T))


def test_round_trip_input_and_uniform_moduli_single_pass():
    # With uniform moduli the wrapper's fixed point is the C0^-1(C0:eps0) round trip
    # (STZ/elastic.c:999-1009): the solve's input differs from eps0 only at ULP level.
    o = _oracle(2, (8, 8, 1))
    rng = np.random.default_rng(9)
    o.field(po.F_EPS0)[:] = rng.standard_normal(3 * o.N) * 0.05
    o.set_scalars(np.zeros(6), np.zeros(6), 0.0)
    o.solve()
    sig1 = o.field(po.F_SIGMA).copy()
    o.solve()
    assert np.array_equal(sig1, o.field(po.F_SIGMA))   # deterministic
