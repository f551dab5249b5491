"""GPU: checkpoint and restart (stz_checkpoint_save / stz_checkpoint_load).

A context restored from a checkpoint continues exactly as the saving context: every
step record (kinds, events, clock, mean strain and stress) and every field bit for bit,
through the device-resident loop and through the host-driven step, in FP64 and mixed
precision, with the full FFT solve and with the event-local update.  Mismatched or
corrupt files are refused with the argument error.
"""
import numpy as np
import pytest

import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

FIELDS = (mg.F_EPS0, mg.F_SIGMA, mg.F_Q0, mg.F_DEPS, mg.F_DEPS_LAST, mg.F_PERM, mg.F_TEMP, mg.F_TLAST)


def _same(a, b, i):
    assert (a.kind, a.site, a.mode, a.n_transformed, a.n_negative, a.printed) == \
        (b.kind, b.site, b.mode, b.n_transformed, b.n_negative, b.printed), i
    assert a.q_chosen == b.q_chosen and a.total_rate == b.total_rate and a.totime == b.totime, i
    assert list(a.epsavg) == list(b.epsavg) and list(a.stress) == list(b.stress), i


def _steps(g, n, resident):
    if resident:
        return g.run_steps(n)[1]
    out = []
    for _ in range(n):
        rc, r = g.step()
        out.append(r)
        if rc:
            break
    return out


CASES = [
    ("fp64_fft_resident", mg.FP64, mg.SOLVER_FFT, True),
    ("fp64_local_resident", mg.FP64, mg.SOLVER_LOCAL, True),
    ("mixed_local_resident", mg.MIXED, mg.SOLVER_LOCAL, True),
    ("fp64_host", mg.FP64, mg.SOLVER_FFT, False),
]


@pytest.mark.parametrize("name,precision,solver,resident", CASES, ids=[c[0] for c in CASES])
def test_restart_continues_exactly(tmp_path, name, precision, solver, resident):
    print("checkpoint", name, flush=True)
    kw = mg.config_3d(n=(16, 16, 16), seed=7, q0_std=0.25, precision=precision)
    a = mg.Stz(**kw)
    a.set_print(5, 1.0)
    a.set_solver(solver, 64, 16)
    _steps(a, 150, resident)
    path = str(tmp_path / "run.ckpt")
    a.checkpoint_save(path)
    ra = _steps(a, 250, resident)
    b = mg.Stz(init_state=False, **dict(kw, seed=99))      # different seed: everything comes from the file
    b.set_solver(solver, 64, 16)
    b.checkpoint_load(path)
    rb = _steps(b, 250, resident)
    assert len(ra) == len(rb) == 250
    assert sum(r.kind == mg.THERMAL for r in ra) > 20
    for i, (x, y) in enumerate(zip(ra, rb)):
        _same(x, y, i)
    for f in FIELDS:
        assert np.array_equal(a.download(f), b.download(f)), f
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert np.array_equal(a.counters(), b.counters())


def test_load_refuses_mismatched_or_corrupt_files(tmp_path):
    kw = mg.config_3d(n=(16, 16, 16), seed=7)
    a = mg.Stz(**kw)
    path = str(tmp_path / "a.ckpt")
    a.checkpoint_save(path)
    other = mg.Stz(init_state=False, **mg.config_3d(n=(16, 16, 8), seed=7))
    with pytest.raises(mg.StzError):
        other.checkpoint_load(path)
    mixed = mg.Stz(init_state=False, **dict(kw, precision=mg.MIXED))
    with pytest.raises(mg.StzError):
        mixed.checkpoint_load(path)
    data = open(path, "rb").read()
    cut = str(tmp_path / "cut.ckpt")
    open(cut, "wb").write(data[: len(data) // 2])
    b = mg.Stz(init_state=False, **kw)
    with pytest.raises(mg.StzError):
        b.checkpoint_load(cut)
    junk = str(tmp_path / "junk.ckpt")
    open(junk, "wb").write(b"not a checkpoint" * 10)
    with pytest.raises(mg.StzError):
        b.checkpoint_load(junk)
    b.checkpoint_load(path)                              # and a good file still loads
    assert np.array_equal(a.download(mg.F_DEPS), b.download(mg.F_DEPS))
