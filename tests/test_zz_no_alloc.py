"""GPU: stz_rates_noalloc (no context, no allocation) against stz_rates on the same fields.

The fields of a context after a few hundred steps (transforms, softening, loading) are
copied into device buffers the test allocates itself; the no-allocation entry point must
give R and the Q<0 count bit for bit equal to the context's rate pass.  Two calls in a row, the
second on a smaller grid (the fixed scratch keeps the first call's contents), plus the
size limit's error.
"""
import ctypes as C

import numpy as np
import pytest

import metallic_glasses_amd as mg

pytestmark = pytest.mark.gpu

FIELDS = [("sigma", mg.F_SIGMA), ("q0", mg.F_Q0), ("deps", mg.F_DEPS), ("deps_last", mg.F_DEPS_LAST),
          ("perm", mg.F_PERM), ("temp", mg.F_TEMP), ("tlast", mg.F_TLAST)]
STZ_E_TOO_LARGE = -6


class _Hip:
    """The test's own device buffers, through the HIP runtime the library already loaded."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0       # host to device
        return p.value

    def get(self, ptr, dtype):
        out = np.zeros(1, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0   # device to host
        return out[0]

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _case(kw, steps):
    g = mg.Stz(**kw)
    g.set_print(5, 1.0)
    rc, recs = g.run_steps(steps)
    assert sum(r.kind == 1 for r in recs) > 0
    nneg_ref, R_ref = g.rates()
    _, _, totime = g.scalars()
    hip = _Hip()
    ptrs = {k: hip.put(g.download(f)) for k, f in FIELDS}
    out_r = hip.put(np.full(1, -1.0))
    out_n = hip.put(np.full(1, -1, dtype=np.int64))
    rc = mg.rates_noalloc(kw, ptrs, totime, out_r, out_n)
    assert rc == 0
    R, n = float(hip.get(out_r, np.float64)), int(hip.get(out_n, np.int64))
    hip.free()
    print(f"  {kw['n']}: R {R!r} (context {R_ref!r}), Q<0 {n} (context {nneg_ref})", flush=True)
    assert R == R_ref
    assert n == nneg_ref
    g.close()


def test_rates_noalloc_equals_context_rates_twice():
    print("rates_noalloc: 3-D 32x32x64 then 2-D 64x64", flush=True)
    _case(mg.config_3d(n=(32, 32, 64), seed=5), 300)
    kw2 = dict(mg.REFERENCE_2D, n=(64, 64, 1), seed=9, loading=mg.SHEAR)
    _case(kw2, 200)


def test_rates_noalloc_rejects_oversized_grids():
    # 20 modes * 2^17 blocks of 256 voxels > 2^21 block sums: refused before any launch
    kw = mg.config_3d(n=(512, 256, 256), seed=1)
    ptrs = {k: 256 for k, _ in FIELDS}          # never dereferenced: the size check comes first
    assert mg.rates_noalloc(kw, ptrs, 0.0, 256, 256) == STZ_E_TOO_LARGE
