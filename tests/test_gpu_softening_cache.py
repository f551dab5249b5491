"""GPU: the cached softening of the rate pass against the oracle when the clock moves back.

A cache entry whose temporary softening term Temp*exp((tlast - t)/tTemp) was exactly zero
is kept as settled and its Perm/Temp/tlast are not read again, which holds only while
the clock advances.  Moving the clock back to just after the latest transform revives
that voxel's temporary term: the rates and the per-voxel minimum Q must then equal the
oracle's for the same state (FP64), and the revived voxel's minimum Q must differ from
its value before the reset (the case a stale settled entry would get wrong).
"""
import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(32, 32), (16, 16, 16)])
def test_clock_reset_revives_temporary_softening(dims):
    print("softening cache", dims, flush=True)
    if len(dims) == 2:
        kw = dict(dim=2, n=(dims[0], dims[1], 1), seed=31, loading=1)
    else:
        kw = dict(dim=3, n=dims, E=94.5, nu=0.35, double_nu=0, gamma0=0.1, seed=31, loading=1)
    o = po.Oracle(**kw)
    g = mg.Stz(**kw)
    o.set_print(5, 1.0)
    g.set_print(5, 1.0)
    for i in range(300):
        rc_o, io = o.step()
        rc_g, ig = g.step()
        assert (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed), i
        assert rc_o == rc_g == 0, i

    def minq_oracle():
        return o.field(po.F_Q).reshape(o.M, o.N).min(axis=0)

    def compare():
        o.softening()
        nneg_o = o.act_energy()
        R_o = o.total_rate()
        nneg_g, R_g = g.rates()
        assert nneg_g == nneg_o
        assert abs(R_g - R_o) <= 1e-13 * R_o
        mo, mgq = minq_oracle(), g.download(mg.F_MINQ)
        assert np.max(np.abs(mgq - mo)) <= 1e-12 * np.max(np.abs(mo))
        return mgq

    compare()
    # the clocks agree to rounding (sums of 1/R in different orders): the GPU takes the
    # oracle's transform times and clock, and one rate pass settles its cache entries again
    tlast = o.field(po.F_TLAST).copy()
    assert np.max(np.abs(g.download(mg.F_TLAST) - tlast)) <= 1e-12 * np.max(np.abs(tlast))
    assert np.array_equal(g.download(mg.F_PERM), o.field(po.F_PERM))
    assert np.array_equal(g.download(mg.F_TEMP), o.field(po.F_TEMP))
    g.upload(mg.F_TLAST, tlast)
    e, s, t = o.scalars()
    g.set_scalars(e, s, t)
    before = compare()
    v = int(np.argmax(tlast))
    assert o.field(po.F_TEMP)[v] > 0
    t_back = float(tlast[v]) + 1e-9          # ~0.006 tTemp after the transform
    assert t_back < t
    o.set_scalars(e, s, t_back)
    g.set_scalars(e, s, t_back)
    after = compare()
    assert after[v] != before[v]             # the temporary term is back in S
