"""The oracle's state image: a run saved after k steps and restored into a fresh state
continues exactly as the uninterrupted run (the long golden traces resume from it)."""
import os

import numpy as np
import pytest

import pyoracle as po


@pytest.mark.parametrize("kw", [dict(dim=2, n=(32, 16, 1), seed=5, loading=1),
                                dict(dim=3, n=(8, 8, 16), seed=9, loading=1, q0_std=0.2)])
def test_state_image_resumes_bit_for_bit(tmp_path, kw):
    kw = dict(kw, modes=4)
    a = po.Oracle(**kw)
    a.set_print(3, 1.0)
    for _ in range(40):
        a.step()
    path = os.path.join(tmp_path, "state.img")
    a.save(path)
    ref = []
    for _ in range(60):
        rc, info = a.step()
        ref.append((rc, info.kind, info.site, info.mode, info.n_transformed, info.totime, info.stress[:6],
                    info.printed))
    b = po.Oracle(empty=True, **kw)
    b.load(path)
    for i in range(60):
        rc, info = b.step()
        got = (rc, info.kind, info.site, info.mode, info.n_transformed, info.totime, info.stress[:6], info.printed)
        assert got == ref[i], i
    for f in (po.F_EPS0, po.F_SIGMA, po.F_Q0, po.F_Q1, po.F_DEPS, po.F_TLAST):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.counters(), b.counters())
    assert a.scalars()[2] == b.scalars()[2]


def test_state_image_rejects_other_shape(tmp_path):
    a = po.Oracle(dim=2, n=(16, 16, 1), modes=2, seed=1)
    path = os.path.join(tmp_path, "state.img")
    a.save(path)
    b = po.Oracle(empty=True, dim=2, n=(32, 16, 1), modes=2, seed=1)
    with pytest.raises(RuntimeError):
        b.load(path)
