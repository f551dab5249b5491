"""GPU event traces against the oracle's committed golden traces (tests/golden/trace_*.npz).

Each fixture holds, per main-loop step of the CPU FP64 oracle, the step kind, the
selected (site, mode) of a thermal event, the number of voxels of an athermal
avalanche, the clock and the loaded component of the mean stress.  The FP64 GPU path
replays the same seed and must reproduce the event sequence exactly and the
stress-strain curve within 1e-10 relative; where the trace holds stress-field snapshots
(a fixed sample of voxels, all components), the field within 1e-10 relative.
"""
import ast
import os

import numpy as np
import pytest

import metallic_glasses_amd as mg
from trace_check import check_sigma, sigma_steps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRACES = sorted(f[6:-4] for f in os.listdir(GOLD) if f.startswith("trace_") and f.endswith(".npz"))


@pytest.mark.parametrize("name", TRACES)
def test_event_trace_matches_oracle(name):
    d = np.load(os.path.join(GOLD, f"trace_{name}.npz"))
    kw = dict(ast.literal_eval(str(d["params"])))
    print("trace", name, kw, len(d["kind"]), "steps", flush=True)
    g = mg.Stz(**kw)
    g.set_print(int(d["printperiod"]), float(d["endstrain"]))
    lc = 0 if kw.get("loading", 0) == 0 else (2 if kw["dim"] == 2 else 5)
    worst = 0.0
    n = len(d["kind"])
    snaps = {s: k for k, s in enumerate(sigma_steps(d))}
    for i in range(n):
        rc, r = g.step()
        if i in snaps:
            check_sigma(g, d, snaps[i])
        k = int(d["kind"][i])
        assert r.kind == k, (i, r.kind, k)
        if k == 1:
            assert (r.site, r.mode) == (int(d["site"][i]), int(d["mode"][i])), i
        elif k == 2:
            assert r.n_transformed == int(d["ntrans"][i]), i
        s = float(d["stress"][i])
        worst = max(worst, abs(r.stress[lc] - s) / max(abs(s), 1e-3))
        assert abs(r.totime - float(d["totime"][i])) <= 1e-10 * max(abs(float(d["totime"][i])), 1e-30), i
    assert worst < 1e-10, worst
    print(f"{name}: {n} steps, {int(np.sum(d['kind'] == 1))} events identical, worst stress {worst:.2e}", flush=True)
