"""Pins of the oracle's FFT (stands in for FFTW 2.x fftwnd_mpi, STZ/init.c:80-83).

Unnormalised, forward sign -1, backward sign +1, as FFTW's FFTW_FORWARD /
FFTW_BACKWARD; the reference divides by N itself (STZ/elastic.c:822).
"""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _dft_bruteforce(a, sign):
    out = a.astype(np.complex128)
    for ax in range(a.ndim):
        n = a.shape[ax]
        k = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("shape", [(4,), (8,), (6,), (12,), (4, 4), (2, 8), (8, 8, 8), (3, 4, 6)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_bruteforce_dft(shape, sign):
    rng = np.random.default_rng(len(shape) * 10 + shape[-1])
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = po.fft3d(a, sign)
    exp = _dft_bruteforce(a, sign)
    assert np.max(np.abs(got - exp)) < 1e-12 * np.max(np.abs(exp))


@pytest.mark.parametrize("shape", [(512, 512), (64, 64, 64), (96, 96), (32, 64, 128)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(3)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = po.fft3d(a, -1)
    exp = np.fft.fftn(a)
    assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-14
    back = po.fft3d(got, +1) / a.size
    assert np.max(np.abs(back - a)) < 1e-13


def test_fft_golden_vectors():
    # fixed vectors written by tests/golden/make_golden.py with numpy.fft
    d = np.load(os.path.join(GOLD, "fft_vectors.npz"))
    for name in d.files:
        if not name.startswith("in_"):
            continue
        key = name[3:]
        a = d[name]
        exp = d["out_" + key]
        got = po.fft3d(a, -1)
        assert np.max(np.abs(got - exp)) <= 1e-13 * np.max(np.abs(exp)), key


def test_fft_parseval_and_linearity():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((16, 16, 16)) + 0j
    b = rng.standard_normal((16, 16, 16)) + 0j
    fa, fb = po.fft3d(a, -1), po.fft3d(b, -1)
    assert abs(np.sum(np.abs(fa) ** 2) / a.size - np.sum(np.abs(a) ** 2)) < 1e-9
    fab = po.fft3d(2.0 * a - 3.0 * b, -1)
    assert np.max(np.abs(fab - (2.0 * fa - 3.0 * fb))) < 1e-11
