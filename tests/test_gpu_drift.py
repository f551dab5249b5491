"""GPU: long-run drift of the mixed-precision path against FP64 (64^3 simple shear, the
event-local update with its periodic full re-solve, 2e4 steps).

Mixed keeps FP64 eigenstrain, clock, mean strain and mean stress, but stores the stress
field in FP32 and updates it event by event, so rounding can build up between full
solves.  Over the whole run: the stress-strain curve agrees with the FP64 run within 1e-4
while the two event sequences are identical, and every 2000 steps the mixed field is
within 1e-4 of the FP64 solve of the mixed run's own eigenstrain field and mean strain
(the same-state error, defined whether or not the trajectories have parted).
"""
import numpy as np
import pytest

import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

STEPS = 20000
CHECK = 2000


def _same_state_err(gm, kw):
    gm.stress_update_local([])          # bring the field to the current mean strain
    f = mg.Stz(init_state=False, **dict(kw, precision=mg.FP64))
    f.upload(mg.F_EPS0, gm.download(mg.F_EPS0))
    e, s, t = gm.scalars()
    f.set_scalars(e, s, t)
    f.stress_solve()
    a = f.download(mg.F_SIGMA)
    b = gm.download(mg.F_SIGMA)
    f.close()
    return float(np.linalg.norm(b - a) / np.linalg.norm(a))


def test_mixed_drift_over_20000_steps():
    kw = mg.config_3d(n=(64, 64, 64), seed=1234, q0_std=0.2)
    runs = {}
    for prec in (mg.FP64, mg.MIXED):
        g = mg.Stz(**dict(kw, precision=prec))
        g.set_print(5, 1.0)
        g.set_solver(mg.SOLVER_LOCAL, 1024, 16)
        runs[prec] = g
    recs = {mg.FP64: [], mg.MIXED: []}
    errs = []
    for chunk in range(STEPS // CHECK):
        for prec, g in runs.items():
            rc, r = g.run_steps(CHECK)
            assert rc >= 0
            recs[prec] += list(r)
        errs.append(_same_state_err(runs[mg.MIXED], kw))
        print(f"  step {(chunk + 1) * CHECK}: same-state error {errs[-1]:.2e}", flush=True)
    a, b = recs[mg.FP64], recs[mg.MIXED]
    same = 0
    for x, y in zip(a, b):
        if (x.kind, x.n_transformed) != (y.kind, y.n_transformed) or \
                (x.kind == 1 and (x.site, x.mode) != (y.site, y.mode)):
            break
        same += 1
    s64 = np.array([r.stress[5] for r in a[:same]])
    smx = np.array([r.stress[5] for r in b[:same]])
    curve = float(np.max(np.abs(smx - s64)) / np.max(np.abs(s64))) if same else 0.0
    events = sum(r.kind == 1 for r in b)
    print(f"  {len(b)} steps ({events} events), identical for {same}, curve error {curve:.2e}, "
          f"worst same-state error {max(errs):.2e}", flush=True)
    assert len(a) == len(b) == STEPS
    assert same > 100
    assert curve < 1e-4
    assert max(errs) < 1e-4
