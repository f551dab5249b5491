"""GPU: dense athermal avalanches against the oracle.

Barriers lowered to -1 eV in 350 voxels (200 of them in the first 256-voxel block) make
the next step an avalanche with many transformed voxels per block, so the workgroup-wide
redraw of a block runs several chunks (more than 20 voxels of M*S values each).  FP64:
the step record and every per-voxel field after the step equal the oracle's bit for bit
(the stress within 1e-10), for glass and crystal-like (MRO) sites, through the per-call
host step and the resident step; the following steps keep the same sequence.
"""
import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

# Tension: the lowered voxels stay below zero and avalanche at every step, and on the
# print steps between them every mean-stress component but xx is at rounding level
# (~1e-17 GPa), where the reference's shear termination test |<sigma_xy>| < 1e-17
# compares rounding noise that no two implementations reproduce alike.
CASES = [
    ("3d_glass", dict(dim=3, n=(16, 16, 16), E=94.5, nu=0.35, double_nu=0, gamma0=0.1, seed=5, loading=0)),
    ("2d_mro", dict(dim=2, n=(32, 32, 1), modes=20, seed=21, loading=0, mro_fraction=0.3, mro_q0=1.5,
                    random_ratio=0.2)),
]

FIELDS = [(mg.F_DEPS, po.F_DEPS), (mg.F_Q0, po.F_Q0), (mg.F_EPS0, po.F_EPS0), (mg.F_PERM, po.F_PERM),
          (mg.F_TEMP, po.F_TEMP), (mg.F_TLAST, po.F_TLAST), (mg.F_DEPS_LAST, po.F_DEPS_LAST)]


def _lowered(kw):
    o = po.Oracle(**kw)
    g = mg.Stz(**kw)
    rng = np.random.default_rng(3)
    idx = np.concatenate([np.arange(0, 200), rng.choice(np.arange(256, o.N), 150, replace=False)])
    q = o.field(po.F_Q0).reshape(o.M, o.N).copy()
    q[:, idx] = -1.0
    o.set_barriers(q.ravel())          # Q0 and the barriers a print step restores
    g.upload(mg.F_Q0, o.field(po.F_Q0))
    o.set_print(5, 1.0)
    g.set_print(5, 1.0)
    return o, g


@pytest.mark.parametrize("resident", [False, True], ids=["host_step", "resident"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_dense_avalanche_matches_oracle(name, kw, resident):
    print("dense avalanche", name, "resident" if resident else "host", flush=True)
    o, g = _lowered(kw)
    rc_o, io = o.step()
    if resident:
        rc_g, recs = g.run_steps(1)
        ig = recs[0]
    else:
        rc_g, ig = g.step()
    assert (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed)
    assert io.n_transformed >= 350
    assert rc_o == rc_g
    for fg, fo in FIELDS:
        assert np.array_equal(g.download(fg), o.field(fo)), fg
    ref = o.field(po.F_SIGMA)
    assert np.max(np.abs(g.download(mg.F_SIGMA) - ref)) <= 1e-10 * np.max(np.abs(ref))
    if name == "2d_mro":
        t = o.mro_map()
        assert np.count_nonzero(t[:200]) > 20      # crystal-like sites took part
    # the run goes on with the same sequence
    for i in range(40):
        rc_o, io = o.step()
        rc_g, ig = g.step()
        assert (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed), i
        if io.kind == 1:
            assert (io.site, io.mode) == (ig.site, ig.mode), i
        assert rc_o == rc_g
        if rc_o:
            break
    assert np.array_equal(g.download(mg.F_DEPS), o.field(po.F_DEPS))
