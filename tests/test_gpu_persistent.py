"""GPU: the persistent step (one kernel launch per kMC step for every replica) against the
kernel sequence (seven launches per shape class and step).

Both run the same item code, so every step record (kind, site, mode, avalanche size, Q<0
count, chosen Q, total rate, clock, mean strain and stress, print flag, |DWork| sum) and
every per-voxel field at the end must be equal bit for bit: single contexts in 2-D and 3-D,
FP64 and mixed, launched directly and replayed from hipGraphs; avalanches of up to 40
voxels taken by the local update on the device; a batched ensemble of uneven grids.  Plans
the persistent step does not take are refused.
"""
import numpy as np
import pytest

import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

FIELDS = [mg.F_SIGMA, mg.F_EPS0, mg.F_Q0, mg.F_DEPS, mg.F_DEPS_LAST, mg.F_PERM, mg.F_TEMP, mg.F_TLAST]


def _rec(r):
    return (r.kind, r.site, r.mode, r.n_transformed, r.n_negative, r.q_chosen, r.total_rate, r.dt_kmc, r.totime,
            tuple(r.stress), tuple(r.epsavg), r.printed, r.soc_elast)


def _ctx(kw, persistent, use_graph=0, resync=1024, local_max=16):
    g = mg.Stz(**kw)
    g.set_print(5, 1.0)
    g.set_solver(mg.SOLVER_LOCAL, resync, local_max)
    g.run_config(use_graph=use_graph)
    if persistent:
        g.persistent(True)
    return g


def _same(a, b, ra, rb, tag):
    assert len(ra) == len(rb), tag
    for i, (x, y) in enumerate(zip(ra, rb)):
        assert _rec(x) == _rec(y), (tag, i)
    for f in FIELDS:
        assert np.array_equal(a.download(f), b.download(f)), (tag, f)


CASES = [
    ("3d_fp64", mg.config_3d(n=(32, 32, 32), seed=5)),
    ("3d_mixed", mg.config_3d(n=(32, 32, 64), seed=6, precision=mg.MIXED)),
    ("3d_mixed_t", mg.config_3d(n=(16, 32, 64), seed=7, precision=mg.MIXED, temperature=420.0)),
    ("2d_fp64", dict(dim=2, n=(128, 256, 1), seed=11)),
    ("2d_mixed_shear", dict(dim=2, n=(256, 128, 1), seed=12, loading=mg.SHEAR, precision=mg.MIXED)),
]


@pytest.mark.parametrize("use_graph", [0, 4], ids=["direct", "graph4"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_persistent_step_equals_kernel_sequence(name, kw, use_graph):
    print("persistent", name, use_graph, flush=True)
    a = _ctx(kw, False)
    b = _ctx(kw, True, use_graph=use_graph)
    rca, ra = a.run_steps(400)
    rcb, rb = b.run_steps(400)
    assert rca == rcb
    _same(a, b, ra, rb, name)
    kinds = [r.kind for r in ra]
    assert kinds.count(mg.THERMAL) > 20
    print(f"{name}: {kinds.count(mg.THERMAL)} thermal, {kinds.count(mg.ATHERMAL)} athermal steps identical",
          flush=True)


def test_persistent_step_device_avalanches():
    """barriers lowered in 40 voxels: avalanches the local update takes on the device
    (local_max 64), through the athermal phase of the persistent step"""
    kw = mg.config_3d(n=(32, 32, 32), seed=9)
    a = _ctx(kw, False, local_max=64)
    b = _ctx(kw, True, local_max=64)
    _, ra0 = a.run_steps(60)
    _, rb0 = b.run_steps(60)
    _same(a, b, ra0, rb0, "warm")
    q = a.download(mg.F_Q0).reshape(a.M, a.N)
    idx = np.random.default_rng(1).choice(a.N, 40, replace=False)
    q[:, idx] = -1.0
    a.upload(mg.F_Q0, q.ravel())
    b.upload(mg.F_Q0, q.ravel())
    _, ra = a.run_steps(200)
    _, rb = b.run_steps(200)
    _same(a, b, ra, rb, "avalanche")
    sizes = [r.n_transformed for r in ra if r.kind == mg.ATHERMAL]
    assert max(sizes) >= 10, sizes
    print("avalanche sizes", sorted(sizes)[-5:], flush=True)


def test_persistent_ensemble_equals_kernel_sequence():
    spec = [((16, 16, 16), 0, 300.0, 0.2), ((32, 32, 32), 1, 400.0, 0.15), ((16, 32, 64), 2, 260.0, 0.3),
            ((16, 16, 16), 3, 350.0, 0.25), ((32, 32, 32), 4, 280.0, 0.1), ((16, 32, 64), 5, 440.0, 0.2),
            ((64, 64, 64), 6, 300.0, 0.2)]
    ens = []
    for persistent in (False, True):
        members = []
        for n, seed, T, sd in spec:
            kw = mg.config_3d(n=n, seed=seed, q0_std=sd, temperature=T, precision=mg.MIXED)
            members.append(_ctx(kw, False))
        e = mg.Ensemble(members, batch=64, use_graph=4 if persistent else 0, persistent=persistent)
        rcs, outs = e.run(300)
        ens.append((members, rcs, outs, e))
    (ma, rca, oa, _), (mb, rcb, ob, _) = ens
    assert rca == rcb
    for i in range(len(spec)):
        _same(ma[i], mb[i], oa[i], ob[i], f"replica {i}")


def test_persistent_refuses_unsupported_plans():
    fft = mg.Stz(**mg.config_3d(n=(16, 16, 16), seed=1))
    with pytest.raises(mg.StzError):
        fft.persistent(True)                    # full FFT solve after every transform
    a = _ctx(mg.config_3d(n=(16, 16, 16), seed=2), False)
    b = _ctx(mg.config_3d(n=(16, 16, 16), seed=3, precision=mg.MIXED), False)
    e = mg.Ensemble([a, b])
    with pytest.raises(mg.StzError):
        e.persistent(True)                      # two precisions in one plan
