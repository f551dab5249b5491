"""Writes the golden fixtures under tests/golden/.

fft_vectors.npz      numpy.fft.fftn of fixed seeded inputs (2-D and 3-D, radix 2/3).
trace_<name>.npz     event traces of the CPU FP64 oracle (oracle/liboracle.so) for the
                     workloads below: per step kind, site, mode, transformed count,
                     clock and loaded-component mean stress.  The GPU parity tests
                     replay the same seed and compare step by step.

Run: python tests/golden/make_golden.py fft
     python tests/golden/make_golden.py trace ref2d_512 [events] [threads] [state_dir]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# reference input file (STZ/input) on the reference's 512x512 2-D grid, tension
REF2D = dict(dim=2, n=(512, 512, 1), seed=123, loading=0)
# 3-D simple shear, mu = 35 GPa, nu = 0.35, gamma0 = 0.1, Q0 ~ N(2.0, 0.2), seed 1234
SHEAR3D = dict(dim=3, E=94.5, nu=0.35, double_nu=0, gamma0=0.1, q0_std=0.2, seed=1234, loading=1,
               athermal_correction=1)
TRACES = {
    "ref2d_512": dict(REF2D),
    "shear3d_64": dict(SHEAR3D, n=(64, 64, 64)),
    "shear3d_128": dict(SHEAR3D, n=(128, 128, 128)),
    "shear2d_256": dict(dim=2, n=(256, 256, 1), seed=7, loading=1),
}


def make_fft():
    rng = np.random.default_rng(20240601)
    out = {}
    for shape in [(8, 8), (16, 32), (4, 8, 16), (6, 12), (12, 6, 24)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = "x".join(str(s) for s in shape)
        out["in_" + key] = a
        out["out_" + key] = np.fft.fftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


# sigma snapshots: the full stress tensor at a fixed seeded sample of voxels, taken after
# the step that completes the given thermal event (print re-solve included)
SIGMA_SAMPLE = 4096
SIGMA_AT = {"shear3d_128": (500, 2000, 10000)}


def sample_voxels(N, k=SIGMA_SAMPLE, seed=20261204):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(N, size=min(k, N), replace=False)).astype(np.int64)


def make_trace(name, events, threads=8, printperiod=5, endstrain=1.0, save_every=1000, state_dir=None):
    """Runs the oracle until `events` thermal events and writes trace_<name>.npz.

    The file is rewritten every `save_every` thermal events, so a long run leaves a
    usable (shorter) trace if it is stopped early.  With `state_dir`, the oracle's whole
    state is written there at the same points, and a later call resumes from the last
    one whose step count matches the trace file."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    po.set_threads(threads)
    kw = TRACES[name]
    t0 = time.time()
    o = po.Oracle(**kw)
    o.set_print(printperiod, endstrain)
    rows = []
    nev = 0
    lc = 0 if kw.get("loading", 0) == 0 else (2 if kw["dim"] == 2 else 5)
    sigma_at = SIGMA_AT.get(name, ())
    vox = sample_voxels(o.N) if sigma_at else np.zeros(0, np.int64)
    snap_step, snap_sigma = [], []
    out = os.path.join(HERE, f"trace_{name}.npz")

    def image(nsteps):
        return os.path.join(state_dir, f"state_{name}_{nsteps}.img")

    if state_dir and os.path.exists(out):
        with np.load(out) as z:
            n = len(z["kind"])
            if os.path.exists(image(n)) and str(z["params"]) == repr(sorted(kw.items())):
                o.load(image(n))
                rows = [tuple(r) for r in zip(z["kind"].tolist(), z["site"].tolist(), z["mode"].tolist(),
                                              z["ntrans"].tolist(), z["totime"].tolist(), z["stress"].tolist(),
                                              z["printed"].tolist())]
                nev = int(np.sum(z["kind"] == 1))
                if sigma_at:
                    snap_step = z["sigma_step"].tolist()
                    snap_sigma = [a for a in z["sigma"]]
                print(f"{name}: resumed at {n} steps, {nev} events", flush=True)

    def save():
        a = np.array(rows, dtype=object)
        tmp = os.path.join(HERE, f".tmp_trace_{name}.npz")
        extra = {}
        if sigma_at:
            extra = dict(sigma_voxels=vox, sigma_step=np.array(snap_step, dtype=np.int64),
                         sigma=np.array(snap_sigma, dtype=np.float64).reshape(len(snap_step), o.C, len(vox)))
        if state_dir:
            os.makedirs(state_dir, exist_ok=True)
            o.save(image(len(rows)))
        np.savez_compressed(tmp,
                            kind=a[:, 0].astype(np.int8), site=a[:, 1].astype(np.int32),
                            mode=a[:, 2].astype(np.int8), ntrans=a[:, 3].astype(np.int32),
                            totime=a[:, 4].astype(np.float64), stress=a[:, 5].astype(np.float64),
                            printed=a[:, 6].astype(np.int8), printperiod=printperiod, endstrain=endstrain,
                            params=np.array(repr(sorted(kw.items()))), **extra)
        os.replace(tmp, out)
        if state_dir:
            for f in os.listdir(state_dir):
                if f.startswith(f"state_{name}_") and f != os.path.basename(image(len(rows))):
                    os.remove(os.path.join(state_dir, f))

    while nev < events:
        rc, info = o.step()
        rows.append((info.kind, info.site, info.mode, info.n_transformed, info.totime, info.stress[lc],
                     info.printed))
        if info.kind == 1:
            nev += 1
            if nev in sigma_at:
                sig = o.field(po.F_SIGMA).reshape(o.C, o.N)
                snap_step.append(len(rows) - 1)
                snap_sigma.append(sig[:, vox].copy())
            if nev % save_every == 0:
                save()
        if len(rows) % 500 == 0:
            print(f"{name}: {len(rows)} steps, {nev} events, {time.time() - t0:.0f}s", flush=True)
        if rc:
            break
    save()
    print(f"{name}: done, {len(rows)} steps, {nev} events, {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    what = sys.argv[1:] or ["fft"]
    if what[0] == "fft":
        make_fft()
    elif what[0] == "trace":
        make_trace(what[1], int(what[2]) if len(what) > 2 else 10000,
                   threads=int(what[3]) if len(what) > 3 else 8,
                   state_dir=what[4] if len(what) > 4 else None)
