"""Builds libstz.so (HIP kernels + C-ABI) and the stz_run driver for gfx950 with hipcc.

Usage: python -m metallic_glasses_amd.build   (or build() from Python)
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libstz.so")
DRIVER = os.path.join(PKG, "stz_run")

ARCH = os.environ.get("STZ_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed FP32 instructions off (wrong results seen with some of their forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function",
]
LIB_SOURCES = ["stz_solve_f64.hip", "stz_solve_f32.hip", "stz_api.hip", "stz_solve.hip", "stz_rates.hip", "stz_resident.hip", "stz_init.hip", "stz_local.hip", "stz_checkpoint.hip", "stz_inhom.hip", "stz_heat.hip", "stz_noalloc.hip", "stz_rng.cpp"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose=False):
    hipcc = _hipcc()
    objdir = os.path.join(PKG, "_objs")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in LIB_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC, "-c",
                   os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out + err)
        elif verbose and (out or err):
            sys.stderr.write(err)
    if failed:
        raise RuntimeError("HIP compilation failed")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB] + objs)
    drv = os.path.join(CSRC, "stz_driver.cpp")
    if os.path.exists(drv):
        _run([hipcc, "-O2", "-std=c++17", "-I", INCLUDE, drv, "-o", DRIVER,
              f"-Wl,-rpath,{PKG}", "-L", PKG, "-lstz"])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
