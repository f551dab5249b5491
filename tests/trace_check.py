"""Helpers of the golden-trace replays: the stress-field snapshots a trace may hold.

A trace written by tests/golden/make_golden.py can carry `sigma_step` (step indices),
`sigma_voxels` (a fixed sample of voxel indices) and `sigma[k][C][len(voxels)]`: the
oracle's full stress tensor at those voxels after step sigma_step[k] (print re-solve
included).
"""
import numpy as np

import metallic_glasses_amd as mg

SIGMA_TOL = 1e-10      # relative, over the sample: |s_gpu - s_oracle|_2 / |s_oracle|_2


def sigma_steps(d):
    return [int(x) for x in d["sigma_step"]] if "sigma_step" in d.files else []


def check_sigma(g, d, k):
    """Compares the context's stress field with snapshot k of the trace; returns the error."""
    vox = d["sigma_voxels"]
    ref = d["sigma"][k]
    C = ref.shape[0]
    got = g.download(mg.F_SIGMA).reshape(C, -1)[:, vox]
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    assert err < SIGMA_TOL, (k, int(d["sigma_step"][k]), err)
    return err


def run_resident_with_snapshots(g, d, n):
    """Resident steps up to n, stopping after each snapshot step to compare the field;
    returns (records, errors)."""
    recs, errs = [], []
    done = 0
    for k, s in enumerate(sigma_steps(d)):
        if s >= n:
            break
        rc, r = g.run_steps(s + 1 - done)
        recs += list(r)
        done += len(r)
        assert done == s + 1, (done, s)
        errs.append(check_sigma(g, d, k))
    if done < n:
        rc, r = g.run_steps(n - done)
        recs += list(r)
    return recs, errs
