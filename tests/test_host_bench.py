"""CPU: the measurement plumbing of bench.py (no GPU, no library calls).

The one-GPU guard, the filter that keeps the dispatches which did a kernel's work, the
64-replica ensemble plan, the comparison of two step sequences, the HBM traffic read
from a committed counter summary, and the rocprofv3 trace summary.
"""
import io
import os
import subprocess
import sys
from contextlib import redirect_stdout

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_active_keeps_dispatches_above_five_percent_of_the_longest():
    v = bench._active([0.001, 10.0, 9.8, 0.2, 0.6])
    assert sorted(v.tolist()) == [0.6, 9.8, 10.0]
    assert bench._active([]).size == 0
    assert bench._active([0.0, 0.0]).tolist() == [0.0, 0.0]


def test_ensemble_plan_is_the_seeded_draw():
    plan = bench.ensemble_plan(64)
    assert len(plan) == 64
    rs = np.random.RandomState(4242)
    for r, (grid, seed, T, sd) in enumerate(plan):
        assert grid == bench.ENSEMBLE_GRIDS[r % 4]
        assert seed == r
        assert T == rs.uniform(250.0, 450.0)
        assert sd == rs.uniform(0.1, 0.3)
        assert 250.0 <= T <= 450.0 and 0.1 <= sd <= 0.3
    assert sorted({p[0] for p in plan}) == sorted(set(bench.ENSEMBLE_GRIDS))
    assert bench.ensemble_plan(64) == plan


def test_step_sequences_compare_on_their_common_prefix():
    # (kind, site, mode, n_transformed, strain, stress)
    a = [(1, 5, 2, 0, 1e-6, 0.10), (0, -1, -1, 0, 2e-6, 0.20), (1, 7, 0, 0, 3e-6, 0.30), (1, 8, 1, 0, 4e-6, 0.4)]
    b = [(1, 5, 2, 0, 1e-6, 0.10), (0, -1, -1, 0, 2e-6, 0.2002), (1, 7, 0, 0, 3e-6, 0.30), (1, 9, 1, 0, 4e-6, 0.4)]
    n = bench.same_prefix(a, b)
    assert n == 3                              # parts at the fourth step (another site)
    assert abs(bench.curve_err(a, b, n) - 0.0002 / 0.30) < 1e-12
    assert bench.curve_err(a, b, 0) is None
    assert bench.same_prefix(a, a[:2]) == 2


def test_traffic_is_read_from_the_committed_counter_summary():
    t = bench.traffic_from_profile((512, 512, 256), "mixed")
    assert t is not None and 3.0e10 < t < 3.7e10      # about 0.94 of 548 B x 67.1 M voxels
    assert bench.traffic_from_profile((512, 512, 256), "fp64") is None
    assert bench.traffic_from_profile((7, 7, 7), "mixed") is None


def test_trace_summary_means_over_all_and_active_dispatches(tmp_path):
    rows = ["Kernel_Name,Start_Timestamp,End_Timestamp"]
    for i, us in enumerate([100.0, 100.0, 1.0, 98.0]):           # one launch returned at once
        rows.append(f"kb_rates,{i * 1000000},{i * 1000000 + us * 1000}")
    rows.append("kb_group,0,20000")
    d = tmp_path / "trace"
    d.mkdir()
    (d / "x_kernel_trace.csv").write_text("\n".join(rows) + "\n")
    buf = io.StringIO()
    with redirect_stdout(buf):
        bench.trace_summary(str(d))
    lines = {ln.split()[0]: ln.split() for ln in buf.getvalue().splitlines()[1:]}
    calls, mean, active, act_mean = lines["kb_rates"][1:5]
    assert int(calls) == 4 and int(active) == 3
    assert abs(float(mean) - 74.75) < 1e-3
    assert abs(float(act_mean) - 298.0 / 3) < 1e-3
    assert int(lines["kb_group"][1]) == 1


def test_physical_cores_counts_core_pairs():
    n = bench.physical_cores()
    assert n is None or 1 <= n <= os.cpu_count()
