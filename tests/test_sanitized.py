"""CPU: host code under AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer.

The library's host random stream: tests/sanitize/rng_check.cpp compiles
metallic_glasses_amd/csrc/stz_rng.cpp (MT19937, skip, the lookup window built on three
threads) and checks it against std::mt19937 and the generator's own Gaussian draws.

The oracle: tests/sanitize/oracle_check.c drives every entry point of oracle/stz_oracle.c on small
2-D and 3-D grids (main-loop steps with print blocks and avalanches, selection, transform,
solve, displacement, heat conduction, per-voxel moduli, crystal-like MRO from a random and
from a given map, a state image written mid-run and continued in a fresh state).  Built
with gcc and the sanitizers (OpenMP on two threads), it must exit 0 with no report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_oracle_runs_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_check")
    build = ["gcc", "-O1", "-g", "-std=gnu99", "-fopenmp", "-ffp-contract=off",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
             "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "stz_oracle.c"),
             os.path.join(ROOT, "tests", "sanitize", "oracle_check.c"), "-lm", "-o", exe]
    b = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(tmp_path / "state.img")], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert r.stdout.startswith("ok"), r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-6000:]
    print(r.stdout.strip())


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_host_random_stream_runs_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "rng_check")
    csrc = os.path.join(ROOT, "metallic_glasses_amd", "csrc")
    build = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-I", csrc,
             os.path.join(csrc, "stz_rng.cpp"), os.path.join(ROOT, "tests", "sanitize", "rng_check.cpp"), "-o", exe]
    b = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert r.stdout.startswith("ok"), r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-6000:]
    print(r.stdout.strip())
