"""Host-side checks of the C-ABI library (no GPU needed).

- libstz.so loads and exports every function include/stz.h declares.
- The MT19937 lookup window of the resident step (StzMtWindow + the kernels' index
  arithmetic) yields exactly the draws of the stream taken in order, and the stream
  itself equals the oracle's pinned generator (GSL mt19937 + polar Gaussian).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "stz.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stz_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(mg.stz.LIB_PATH)
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _oracle_draws(seed, skip, with_uniform, ngauss):
    r = po.OraMT()
    L = po.lib()
    L.ora_mt_seed(C.byref(r), seed)
    for _ in range(skip):
        L.ora_mt_next(C.byref(r))
    out = []
    if with_uniform:
        out.append(L.ora_mt_next(C.byref(r)) / 4294967296.0)
    for _ in range(ngauss):
        out.append(L.ora_gaussian(C.byref(r), 1.0))
    return np.array(out)


CASES = [
    # seed, skip, window, offset, uniform first, gaussians
    (1234, 0, 4096, 0, 1, 100),
    (1234, 0, 4096, 1, 1, 100),
    (1234, 7, 4096, 2, 0, 300),
    (1234, 7, 4096, 3, 0, 300),
    (123, 1000, 1 << 16, 777, 1, 2000),
    (0, 5, 8192, 4000, 0, 1000),        # seed 0 -> 4357
    (4242, 123457, 1 << 15, 12345, 1, 5000),
]


@pytest.mark.parametrize("seed,skip,window,offset,uni,ng", CASES)
def test_window_draws_equal_stream(seed, skip, window, offset, uni, ng):
    w, s, we, se = mg.mt_window_probe(seed, skip, window, offset, uni, ng)
    assert np.array_equal(w, s)
    assert we == se


@pytest.mark.parametrize("seed,skip,offset,uni,ng", [(1234, 0, 0, 1, 50), (99, 3, 10, 0, 200), (7, 1000, 1, 1, 400)])
def test_stream_equals_oracle_generator(seed, skip, offset, uni, ng):
    _, s, _, _ = mg.mt_window_probe(seed, skip, 1 << 14, offset, uni, ng)
    ref = _oracle_draws(seed, skip + offset, uni, ng)
    assert np.array_equal(s, ref)


def test_window_end_positions_consume_two_words_per_attempt():
    # 100 gaussians take ~127 attempts (acceptance pi/4), two words each
    _, _, we, _ = mg.mt_window_probe(5, 0, 4096, 0, 0, 100)
    assert we % 2 == 0 and 200 <= we <= 2 * 200


def test_window_too_short_is_an_error():
    with pytest.raises(mg.StzError):
        mg.mt_window_probe(5, 0, 64, 0, 0, 100)


# Philox4x32-10 known answers (Random123 kat_vectors: philox4x32_10)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    lib = mg.lib()
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    o = (C.c_uint * 4)()
    assert lib.stz_philox_probe(c, k, o) == 0
    assert tuple(o) == expect
