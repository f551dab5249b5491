"""CPU pins of the oracle's print-step displacement (ora_displacement, the restatement of
InhomogeneousStressSolver(uflag=2): STZ/bmg.c:289-295, STZ/elastic.c:873-900).

- Compatibility: the displacement fluctuation's symmetric spectral gradient,
  eps_hat_ab = (i/2)(k_a u_hat_b + k_b u_hat_a), equals the solved strain fluctuation on
  every wavevector with no Nyquist component (k from the reference's grid,
  k = 2*pi*(j/n - round(j/n)), x = the fast axis).  A flipped sign of i, a transposed
  component or a missing 1/|k| fails it.
- Position: with eps0 = 0 the fluctuation vanishes and u is the supercell position,
  which the reference adds with its slow index along x in 2-D
  (uo[0] += i*(1+e_xx)*LX + j*e_xy*LY, STZ/elastic.c:896-899).
"""
import numpy as np
import pytest

import pyoracle as po

COMP = {2: [(0, 0), (1, 1), (0, 1)], 3: [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]}
PI_REF = 3.14159265359


def _kw(dims):
    if len(dims) == 2:
        return dict(dim=2, n=(dims[0], dims[1], 1))
    return dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1)


def _position(o, epsavg):
    dim = o.params.dim
    n0, n1, n2 = o.dims
    L = (o.params.lx, o.params.ly, o.params.lz)
    i2, i1, i0 = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    e = epsavg
    if dim == 2:
        i, j = i1.astype(float), i0.astype(float)
        return np.stack([i * (1 + e[0]) * L[0] + j * e[2] * L[1], i * e[2] * L[0] + j * (1 + e[1]) * L[1]])
    idx = (i0.astype(float), i1.astype(float), i2.astype(float))
    vo = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    out = []
    for a in range(3):
        pos = np.zeros(i0.shape)
        for b in range(3):
            pos = pos + idx[b] * ((1.0 if a == b else 0.0) + e[vo[a][b]]) * L[b]
        out.append(pos)
    return np.stack(out)


GRIDS = [(16, 8), (32, 32), (8, 8, 8), (8, 4, 16)]


@pytest.mark.parametrize("dims", GRIDS)
def test_displacement_gradient_is_the_solved_strain(dims):
    o = po.Oracle(empty=True, **_kw(dims))
    dim = o.params.dim
    rng = np.random.default_rng(3)
    o.field(po.F_EPS0)[:] = rng.standard_normal(o.C * o.N) * 0.05
    ea = rng.standard_normal(6) * 0.01
    o.set_scalars(ea, np.zeros(6), 0.0)
    u = o.displacement()
    n0, n1, n2 = o.dims
    shape = (n2, n1, n0)
    uf = u.reshape(dim, *shape) - _position(o, ea)[:dim]
    epsf = o.field(po.F_EPSO).reshape(o.C, *shape) - ea[: o.C, None, None, None]
    uh = np.fft.fftn(uf, axes=(1, 2, 3))
    eh = np.fft.fftn(epsf, axes=(1, 2, 3))
    ks, nyq = [], np.zeros(shape, dtype=bool)
    for a, n in enumerate((n0, n1, n2)):
        j = np.arange(n)
        s = j * (1.0 / n)
        k1 = (s - np.round(s)) * 2 * PI_REF
        kk = np.zeros(shape)
        ny = np.zeros(shape, dtype=bool)
        sl = [None, None, None]
        sl[2 - a] = slice(None)
        kk = kk + k1[tuple(sl)]
        if n > 1:
            ny = ny | (j == n // 2)[tuple(sl)]
        ks.append(kk)
        nyq |= ny
    mask = ~nyq
    scale = np.max(np.abs(eh))
    for c, (a, b) in enumerate(COMP[dim]):
        pred = 0.5j * (ks[a] * uh[b] + ks[b] * uh[a])
        assert np.max(np.abs((pred - eh[c])[mask])) <= 1e-11 * scale, (c, a, b)
    assert mask.sum() > 0.5 * mask.size


@pytest.mark.parametrize("dims", GRIDS)
def test_displacement_without_eigenstrain_is_the_position(dims):
    o = po.Oracle(empty=True, **_kw(dims))
    dim = o.params.dim
    ea = np.array([0.01, -0.004, 0.002, 0.003, -0.001, 0.005])
    o.set_scalars(ea, np.zeros(6), 0.0)
    u = o.displacement()
    n0, n1, n2 = o.dims
    pos = _position(o, ea)[:dim].reshape(dim, -1)
    assert np.max(np.abs(u.reshape(dim, -1) - pos)) <= 1e-12 * np.max(np.abs(pos))
