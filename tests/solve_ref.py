"""Independent numpy formulations used to pin the stress solve.

Nothing here calls the oracle or the product: the Green's operator is applied
by solving the acoustic-tensor system K(n) w = sigma0_hat . n per wave vector
with numpy.linalg.solve, instead of the closed-form alpha expression the
reference uses (STZ/elastic.c:786-809).
"""
import numpy as np

PAIRS2 = [(0, 0), (1, 1), (0, 1)]
PAIRS3 = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
PI_REF = 3.14159265359  # truncated PI of STZ/bmg.h:74


def pairs(dim):
    return PAIRS2 if dim == 2 else PAIRS3


def elastic_constants(E, nu, dim, double_nu):
    if double_nu:
        nu = 2 * nu
    mu = E / (2 + 2 * nu)
    lam = E * nu / (1 - nu * nu) if dim == 2 else E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, mu, nu


def kvec(dims, dim):
    """Per-axis k with the reference convention (s - round_half_away(s)) * 2 PI, STZ/init.c:126-129."""
    ks = []
    for a in range(dim):
        n = dims[a]
        s = np.arange(n) * (1.0 / n)
        ks.append((s - np.floor(s + 0.5)) * 2 * PI_REF)
    return ks


def to_tensor(f, dim):
    """[C, ...] Voigt-like field -> [..., dim, dim] symmetric tensor."""
    P = pairs(dim)
    shp = f.shape[1:]
    t = np.zeros(shp + (dim, dim), dtype=f.dtype)
    for c, (i, j) in enumerate(P):
        t[..., i, j] = f[c]
        t[..., j, i] = f[c]
    return t


def from_tensor(t, dim):
    return np.stack([t[..., i, j] for (i, j) in pairs(dim)])


def c0_mul(t, lam, mu, dim):
    tr = np.trace(t, axis1=-2, axis2=-1)
    return lam * tr[..., None, None] * np.eye(dim) + 2 * mu * t


def green_solve(eps0, epsavg, dims, dim, lam, mu):
    """eps0: [C, n_(dim-1), ..., n_0] real.  Returns (eps_total, sigma) like the reference.

    Full complex spectrum, each bin with its own unit vector (including the
    reference's Nyquist sign), then the real part of the inverse transform.
    """
    axes = tuple(range(1, dim + 1))
    e_hat = np.fft.fftn(eps0, axes=axes)                # [C, ...]
    t_hat = to_tensor(e_hat, dim)                       # [..., d, d]
    s_hat = c0_mul(t_hat, lam, mu, dim)
    ks = kvec(dims, dim)
    # grids in array order (slowest first): array axis order is (n_{dim-1}, ..., n_0)
    grids = np.meshgrid(*[ks[a] for a in reversed(range(dim))], indexing="ij")
    kv = np.stack([grids[dim - 1 - a] for a in range(dim)], axis=-1)   # [..., dim] component a = axis a
    kn = np.linalg.norm(kv, axis=-1)
    zero = kn == 0
    kn[zero] = 1.0
    n = kv / kn[..., None]
    F = np.einsum("...ij,...j->...i", s_hat, n)
    K = mu * np.eye(dim) + (lam + mu) * np.einsum("...i,...j->...ij", n, n)
    K[zero] = np.eye(dim)
    w = np.linalg.solve(K, F[..., None])[..., 0]
    eh = 0.5 * (np.einsum("...i,...j->...ij", n, w) + np.einsum("...i,...j->...ij", w, n))
    eh[zero] = 0
    e_tot_hat = from_tensor(eh, dim)
    eps = np.real(np.fft.ifftn(e_tot_hat, axes=axes)) + np.asarray(epsavg)[: eps0.shape[0]].reshape((-1,) + (1,) * dim)
    sig = from_tensor(c0_mul(to_tensor(eps - eps0, dim), lam, mu, dim), dim)
    return eps, sig
