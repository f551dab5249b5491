"""Writes the small golden fixtures under tests/golden/.

fft_vectors.npz: numpy.fft.fftn of fixed seeded inputs (2-D and 3-D, radix 2/3).
Run: python tests/golden/make_golden.py fft
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def make_fft():
    rng = np.random.default_rng(20240601)
    out = {}
    for shape in [(8, 8), (16, 32), (4, 8, 16), (6, 12), (12, 6, 24)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = "x".join(str(s) for s in shape)
        out["in_" + key] = a
        out["out_" + key] = np.fft.fftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


if __name__ == "__main__":
    what = sys.argv[1:] or ["fft"]
    if "fft" in what:
        make_fft()
