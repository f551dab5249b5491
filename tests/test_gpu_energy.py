"""GPU: ElasticEnergy (STZ/elastic.c:533-547, computed every step at STZ/bmg.c:174-178 and
recorded at print steps, :225) against the oracle, whose value the CPU tests pin against
sigma:C0^-1:sigma/2 with a numerical inverse of the stiffness.

Tolerances: 1e-12 relative in FP64 (the GPU sums per 256-voxel block, then the blocks;
the oracle sums voxel by voxel), 1e-6 in mixed precision (FP32 stress field).
"""
import numpy as np
import pytest

import pyoracle as po
import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

CASES = [
    ("2d_64", dict(dim=2, n=(64, 64, 1), loading=0)),
    ("2d_512", dict(dim=2, n=(512, 512, 1), loading=1)),
    ("3d_16", dict(dim=3, n=(16, 16, 16), E=94.5, nu=0.35, double_nu=0, gamma0=0.1, loading=1,
                   athermal_correction=1)),
    ("3d_32x32x64", dict(dim=3, n=(32, 32, 64), E=94.5, nu=0.35, double_nu=0, gamma0=0.1, loading=0,
                         athermal_correction=1)),
]
TOL = {mg.FP64: 1e-12, mg.MIXED: 1e-6}


@pytest.mark.parametrize("precision", [mg.FP64, mg.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_elastic_energy_of_an_uploaded_field(name, kw, precision):
    """A random stress field, then a pending uniform increment on one component (applied
    lazily in FP64, kept as an offset in mixed precision)."""
    print("energy upload", name, precision, flush=True)
    o = po.Oracle(empty=True, **kw)
    g = mg.Stz(init_state=False, precision=precision, **kw)
    rng = np.random.default_rng(7)
    sig = rng.standard_normal(o.C * o.N) * 0.8
    o.field(po.F_SIGMA)[:] = sig
    g.upload(mg.F_SIGMA, sig)
    e_o = o.elastic_energy()
    assert abs(g.elastic_energy() - e_o) <= TOL[precision] * abs(e_o)
    comp = o.C - 1
    g.add_uniform_stress(comp, 0.37)
    o.field(po.F_SIGMA).reshape(o.C, o.N)[comp] += 0.37
    e_o = o.elastic_energy()
    assert abs(g.elastic_energy() - e_o) <= TOL[precision] * abs(e_o)


@pytest.mark.parametrize("precision", [mg.FP64, mg.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,kw", [CASES[0], CASES[2]], ids=[CASES[0][0], CASES[2][0]])
def test_elastic_energy_at_print_steps(name, kw, precision):
    """The value a print block records: after the step's loading, before the re-solve."""
    print("energy trajectory", name, precision, flush=True)
    o = po.Oracle(**kw)
    g = mg.Stz(precision=precision, **kw)
    o.set_print(1, 1.0)
    g.set_print(1, 1.0)
    printed = 0
    for i in range(300):
        rc_o, io = o.step()
        rc_g, ig = g.step_kmc()
        same = (io.kind, io.n_transformed) == (ig.kind, ig.n_transformed)
        if io.kind == 1:                # site and mode are defined for thermal events only
            same = same and (io.site, io.mode) == (ig.site, ig.mode)
        if precision == mg.MIXED and not same:
            break                       # trajectories may part in mixed precision
        assert same, i
        assert io.printed == ig.printed
        if ig.printed:
            e_g = g.elastic_energy()
            assert abs(e_g - io.el_energy) <= TOL[precision] * abs(io.el_energy), i
            printed += 1
            g.step_print(ig)
        if rc_o:
            break
    assert printed > 20
