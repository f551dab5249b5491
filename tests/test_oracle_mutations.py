"""The step pins catch plausible mistakes: each case plants one mistake in a copy of the
oracle's C source (on the CPU), builds it, and runs tests/test_oracle_step_pins.py against
that build in a child process, which must fail.  The unmodified source must pass the same
way (the harness itself is sound)."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "stz_oracle.c")
CFLAGS = ["-O2", "-std=gnu99", "-fopenmp", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]

# (name, exact text in the oracle, its planted replacement)
MUTANTS = [
    ("athermal correction added instead of subtracted",
     "for (int c = 0; c < s->C; c++) s->epsavg[c] -= te[c];",
     "for (int c = 0; c < s->C; c++) s->epsavg[c] += te[c];"),
    ("shear elastic step adds the strain increment instead of strainrate*1",
     "s->epsavg[lc] += sr * 1;",
     "s->epsavg[lc] += inc;"),
    ("direction modifier linear in D",
     "dm = (1 + dm) * (1 + dm) / 4;",
     "dm = (1 + dm) / 2;"),
    ("tension thermal step drops the athermal term of the transverse strain",
     "s->epsavg[1] += (sr * t_kmc * (-s->nu) - te[1]);",
     "s->epsavg[1] += (sr * t_kmc * (-s->nu));"),
    ("no barrier reset at print steps",
     "memcpy(s->q0, s->q1, sizeof(double) * (size_t)s->M * (size_t)s->N);",
     ";"),
    ("athermal threshold inclusive",
     "fabs(minq) < s->p.athermal_threshold",
     "fabs(minq) <= s->p.athermal_threshold"),
    ("argmin takes the last of equal barriers",
     "if (minq > s->q[(long)m * N + v])",
     "if (minq >= s->q[(long)m * N + v])"),
    ("elastic step advances the clock by the strain increment",
     "s->totime += s->dt;",
     "s->totime += s->p.strainincrement;"),
]


def _build(text, out):
    src = out + ".c"
    with open(src, "w") as f:
        f.write(text)
    subprocess.run(["gcc"] + CFLAGS + ["-I", os.path.join(ROOT, "oracle"), "-shared", "-o", out, src,
                    "-lm", "-lgomp"], check=True, capture_output=True)


def _pins_pass(lib):
    env = dict(os.environ, STZ_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_oracle_step_pins.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode == 0


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


def test_unmodified_copy_passes(source, tmp_path):
    lib = str(tmp_path / "liboracle_copy.so")
    _build(source, lib)
    assert _pins_pass(lib)


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(source, tmp_path, name, old, new):
    assert source.count(old) == 1, f"the text of '{name}' is not in the oracle exactly once"
    lib = str(tmp_path / "liboracle_mutant.so")
    _build(source.replace(old, new), lib)
    assert not _pins_pass(lib), f"the step pins did not catch: {name}"
