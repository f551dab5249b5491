"""GPU: the 64-replica ensemble of uneven grids at its real size.

Replica r: grid r mod 4 of {32^3, 64x64x128, 96^3, 128^3}, seed r, and (T, barrier std)
drawn in pairs from MT19937(4242) (numpy's legacy generator): T ~ U[250, 450] K, std ~
U[0.1, 0.3] eV; 3-D simple shear, the event-local update.  The batched mixed ensemble must
equal every replica stepped alone (records and final stress field bit for bit), and every
mixed replica must follow its FP64 twin (the FP64 ensemble) within 1e-4 on the
stress-strain curve while their event sequences agree, with the same-state field error
below 1e-4 at the end.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32, 32), (64, 64, 128), (96, 96, 96), (128, 128, 128)]
STEPS = 120


def _plan(n=64):
    rs = np.random.RandomState(4242)
    out = []
    for r in range(n):
        T = float(rs.uniform(250.0, 450.0))
        sd = float(rs.uniform(0.1, 0.3))
        out.append(mg.config_3d(n=GRIDS[r % 4], seed=r, q0_std=sd, temperature=T))
    return out


def _make(kws):
    def one(kw):
        g = mg.Stz(**kw)
        g.set_print(5, 1.0)
        g.set_solver(mg.SOLVER_LOCAL, 1024, 16)
        return g
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(one, kws))


def _key(r):
    return (r.kind, r.n_transformed) + ((r.site, r.mode) if r.kind == 1 else ())


def test_64_replicas_equal_single_runs_and_track_fp64():
    kws = _plan()
    mixed_kws = [dict(k, precision=mg.MIXED) for k in kws]
    print("64 replicas: 64 mixed replicas", flush=True)
    members = _make(mixed_kws)
    ens = mg.Ensemble(members, batch=64, use_graph=4)
    rcs, outs = ens.run(STEPS)
    assert all(rc == 0 for rc in rcs) and all(len(o) == STEPS for o in outs)
    print("64 replicas: each replica alone", flush=True)
    for i in range(0, 64, 16):
        singles = _make(mixed_kws[i:i + 16])
        for j, g in enumerate(singles):
            rc, rec = g.run_steps(STEPS)
            assert len(rec) == STEPS
            for s, (a, b) in enumerate(zip(outs[i + j], rec)):
                assert _key(a) == _key(b) and list(a.stress) == list(b.stress), (i + j, s)
            assert np.array_equal(members[i + j].download(mg.F_SIGMA), g.download(mg.F_SIGMA)), i + j
            g.close()
    print("64 replicas: FP64 twin ensemble", flush=True)
    twins = _make(kws)
    tens = mg.Ensemble(twins, batch=64, use_graph=4)
    _, outs64 = tens.run(STEPS)
    worst_curve, worst_state, parted = 0.0, 0.0, 0
    for i in range(64):
        same = 0
        for a, b in zip(outs[i], outs64[i]):
            if _key(a) != _key(b):
                break
            same += 1
        parted += same < STEPS
        if same:
            s64 = np.array([r.stress[5] for r in outs64[i][:same]])
            smx = np.array([r.stress[5] for r in outs[i][:same]])
            worst_curve = max(worst_curve, float(np.max(np.abs(smx - s64)) / np.max(np.abs(s64))))
        # same-state error: the FP64 solve of the mixed replica's eigenstrain and mean strain
        m = members[i]
        m.stress_update_local([])
        f = twins[i]
        f.upload(mg.F_EPS0, m.download(mg.F_EPS0))
        e, s, t = m.scalars()
        f.set_scalars(e, s, t)
        f.stress_solve()
        a, b = f.download(mg.F_SIGMA), m.download(mg.F_SIGMA)
        worst_state = max(worst_state, float(np.linalg.norm(b - a) / np.linalg.norm(a)))
    print(f"64 replicas: {parted} of 64 replicas parted from FP64 within {STEPS} steps; worst curve error "
          f"{worst_curve:.2e}, worst same-state error {worst_state:.2e}", flush=True)
    assert worst_curve < 1e-4
    assert worst_state < 1e-4
    tens.close()
    ens.close()


def test_64_replicas_persistent_step_equals_kernel_sequence():
    """The same 64 mixed replicas as one persistent launch per step (replayed 4 steps per
    graph) against the kernel sequence: every record and every final stress and
    eigenstrain field equal bit for bit."""
    kws = [dict(k, precision=mg.MIXED) for k in _plan()]
    runs = []
    for persistent in (False, True):
        print("64 replicas, persistent" if persistent else "64 replicas, kernel sequence", flush=True)
        members = _make(kws)
        ens = mg.Ensemble(members, batch=64, use_graph=4, persistent=persistent)
        rcs, outs = ens.run(60)
        runs.append((members, rcs, outs, ens))
    (ma, rca, oa, ea), (mb, rcb, ob, eb) = runs
    assert rca == rcb
    for i in range(64):
        assert len(oa[i]) == len(ob[i]) == 60, i
        for s, (x, y) in enumerate(zip(oa[i], ob[i])):
            assert _key(x) == _key(y) and list(x.stress) == list(y.stress) and x.totime == y.totime, (i, s)
        for f in (mg.F_SIGMA, mg.F_EPS0):
            assert np.array_equal(ma[i].download(f), mb[i].download(f)), (i, f)
    ea.close()
    eb.close()
