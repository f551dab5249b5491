"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# STZ_ORACLE_LIB: another build of the oracle (the mutation tests load planted-mistake copies)
LIB_PATH = os.environ.get("STZ_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OraParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n", C.c_int * 3), ("modes", C.c_int), ("precision", C.c_int),
        ("loading", C.c_int), ("athermal_correction", C.c_int), ("softening", C.c_int),
        ("seed", C.c_int),
        ("E", C.c_double), ("nu", C.c_double), ("double_nu", C.c_int),
        ("q0_mean", C.c_double), ("q0_std", C.c_double), ("gamma0", C.c_double),
        ("temperature", C.c_double), ("trial_frequency", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("strainrate", C.c_double), ("strainincrement", C.c_double),
        ("perm_soft", C.c_double), ("temp_soft", C.c_double), ("perm_cap", C.c_double),
        ("temp_soft_barrier", C.c_double), ("athermal_threshold", C.c_double),
        ("mro_fraction", C.c_double), ("mro_q0", C.c_double), ("random_ratio", C.c_double),
    ]


class OraStepInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("site", C.c_int), ("mode", C.c_int), ("n_transformed", C.c_int),
        ("n_negative", C.c_longlong), ("q_chosen", C.c_double), ("total_rate", C.c_double),
        ("dt_kmc", C.c_double), ("totime", C.c_double), ("epsavg", C.c_double * 6),
        ("stress", C.c_double * 6), ("printed", C.c_int), ("mro", C.c_int), ("el_energy", C.c_double),
    ]


class OraMT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


F_EPS0, F_SIGMA, F_Q0, F_DEPS, F_DEPS_LAST, F_PERM, F_TEMP, F_TLAST, F_Q, F_SOFT, F_EPSO, F_TLOCAL, F_Q1 = range(13)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.ora_mt_seed.argtypes = [P(OraMT), C.c_ulong]
        L.ora_mt_next.argtypes = [P(OraMT)]
        L.ora_mt_next.restype = C.c_uint32
        L.ora_mt_uniform.argtypes = [P(OraMT)]
        L.ora_mt_uniform.restype = C.c_double
        L.ora_gaussian.argtypes = [P(OraMT), C.c_double]
        L.ora_gaussian.restype = C.c_double
        L.ora_fft3d.argtypes = [P(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int]
        L.ora_create.argtypes = [P(OraParams)]
        L.ora_create.restype = C.c_void_p
        L.ora_create_empty.argtypes = [P(OraParams)]
        L.ora_create_empty.restype = C.c_void_p
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_field.argtypes = [C.c_void_p, C.c_int]
        L.ora_field.restype = P(C.c_double)
        L.ora_scalars.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ora_set_scalars.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_double]
        L.ora_consts.argtypes = [C.c_void_p, P(C.c_double)]
        L.ora_rng.argtypes = [C.c_void_p]
        L.ora_rng.restype = P(OraMT)
        L.ora_counters.argtypes = [C.c_void_p, P(C.c_longlong)]
        L.ora_stress_solve.argtypes = [C.c_void_p]
        L.ora_softening.argtypes = [C.c_void_p]
        L.ora_act_energy.argtypes = [C.c_void_p]
        L.ora_act_energy.restype = C.c_longlong
        L.ora_total_rate.argtypes = [C.c_void_p]
        L.ora_total_rate.restype = C.c_double
        L.ora_select.argtypes = [C.c_void_p, C.c_double, P(C.c_int), P(C.c_int)]
        L.ora_transform.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ora_elastic_energy.argtypes = [C.c_void_p]
        L.ora_elastic_energy.restype = C.c_double
        L.ora_step.argtypes = [C.c_void_p, P(OraStepInfo)]
        L.ora_set_print.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.ora_set_strainhist.argtypes = [C.c_void_p, C.c_double]
        L.ora_displacement.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ora_set_moduli.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ora_heat_conduct.argtypes = [C.c_void_p, C.c_double]
        L.ora_create_mro.argtypes = [C.POINTER(OraParams), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ora_create_mro.restype = C.c_void_p
        L.ora_mro_map.argtypes = [C.c_void_p]
        L.ora_mro_map.restype = C.POINTER(C.c_ubyte)
        L.ora_set_heat.argtypes = [C.c_void_p, C.c_int]
        L.ora_inhom_iterations.argtypes = [C.c_void_p]
        L.ora_inhom_iterations.restype = C.c_int
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_set_threads.restype = C.c_int
        L.ora_save.argtypes = [C.c_void_p, C.c_char_p]
        L.ora_save.restype = C.c_int
        L.ora_load.argtypes = [C.c_void_p, C.c_char_p]
        L.ora_load.restype = C.c_int
        _lib = L
    return _lib


# default physics of the reference input (STZ/input, STZ/bmg.h, STZ/init.c:60-68)
REFERENCE_2D = dict(
    dim=2, n=(512, 512, 1), modes=20, precision=0, loading=0, athermal_correction=0,
    softening=1, seed=123, E=69.46, nu=0.365, double_nu=1, q0_mean=2.0, q0_std=0.25,
    gamma0=0.14, temperature=300.0, trial_frequency=1e13, lx=0.7, ly=0.7, lz=1.0,
    strainrate=1e-4, strainincrement=1e-4, perm_soft=13.0, temp_soft=39.0,
    perm_cap=float(-np.log(0.6)), temp_soft_barrier=0.37, athermal_threshold=0.33,
    mro_fraction=0.0, mro_q0=2.0, random_ratio=0.2,
)


def make_params(**kw):
    d = dict(REFERENCE_2D)
    d.update(kw)
    p = OraParams()
    for k, v in d.items():
        if k == "n":
            for i in range(3):
                p.n[i] = int(v[i]) if i < len(v) else 1
        else:
            setattr(p, k, v)
    return p


class Oracle:
    """One oracle state (the reference's global state, one rank)."""

    def __init__(self, empty=False, mro_map=None, **kw):
        """mro_map = (types[N], thetas[N]): the InitMRO 1 path (a map read from a file)."""
        self.params = make_params(**kw)
        L = lib()
        if mro_map is not None:
            t = np.ascontiguousarray(mro_map[0], dtype=np.int32)
            th = np.ascontiguousarray(mro_map[1], dtype=np.float64)
            self._h = L.ora_create_mro(C.byref(self.params), t.ctypes.data_as(C.POINTER(C.c_int)),
                                       th.ctypes.data_as(C.POINTER(C.c_double)))
        else:
            self._h = (L.ora_create_empty if empty else L.ora_create)(C.byref(self.params))
        if not self._h:
            raise RuntimeError("ora_create failed")
        c = np.zeros(16)
        L.ora_consts(self._h, c.ctypes.data_as(C.POINTER(C.c_double)))
        self.consts = c
        self.N = int(c[10]); self.C = int(c[11]); self.S = int(c[12]); self.M = int(c[13])
        self.dims = tuple(int(x) for x in self.params.n)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ora_destroy(h)
            self._h = None

    def field(self, which):
        n = {F_EPS0: self.C * self.N, F_SIGMA: self.C * self.N, F_EPSO: self.C * self.N,
             F_Q0: self.M * self.N, F_Q1: self.M * self.N, F_Q: self.M * self.N, F_DEPS: self.M * self.S * self.N,
             F_DEPS_LAST: self.S * self.N}.get(which, self.N)
        ptr = lib().ora_field(self._h, which)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def set_barriers(self, q0):
        """Sets Q0 and the stress-free barriers Q1 a print step restores (as stz_upload of
        STZ_F_Q0 does on the GPU side)."""
        self.field(F_Q0)[:] = q0
        self.field(F_Q1)[:] = q0

    def scalars(self):
        e = np.zeros(6); s = np.zeros(6); t = C.c_double()
        dp = C.POINTER(C.c_double)
        lib().ora_scalars(self._h, e.ctypes.data_as(dp), s.ctypes.data_as(dp), C.byref(t))
        return e, s, t.value

    def set_scalars(self, epsavg, stress, totime):
        dp = C.POINTER(C.c_double)
        e = np.ascontiguousarray(epsavg, dtype=np.float64)
        s = np.ascontiguousarray(stress, dtype=np.float64)
        lib().ora_set_scalars(self._h, e.ctypes.data_as(dp), s.ctypes.data_as(dp), float(totime))

    def rng(self):
        return lib().ora_rng(self._h)

    def counters(self):
        o = np.zeros(8, dtype=np.int64)
        lib().ora_counters(self._h, o.ctypes.data_as(C.POINTER(C.c_longlong)))
        return o

    def solve(self):
        lib().ora_stress_solve(self._h)

    def set_moduli(self, lam, mu):
        """Per-voxel Lambda, Mu (None: uniform).  Keeps references to the arrays' copies."""
        if lam is None:
            lib().ora_set_moduli(self._h, None, None)
            return
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._mu = np.ascontiguousarray(mu, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        lib().ora_set_moduli(self._h, self._lam.ctypes.data_as(dp), self._mu.ctypes.data_as(dp))

    def mro_map(self):
        return np.ctypeslib.as_array(lib().ora_mro_map(self._h), shape=(self.N,)).copy()

    def heat_conduct(self, dt):
        return lib().ora_heat_conduct(self._h, float(dt))

    def set_heat(self, on):
        lib().ora_set_heat(self._h, 1 if on else 0)

    def inhom_iterations(self):
        return lib().ora_inhom_iterations(self._h)

    def displacement(self):
        """Print-step re-solve and displacement field u[dim][N]."""
        u = np.zeros(self.params.dim * self.N)
        lib().ora_displacement(self._h, u.ctypes.data_as(C.POINTER(C.c_double)))
        return u

    def softening(self):
        lib().ora_softening(self._h)

    def act_energy(self):
        return lib().ora_act_energy(self._h)

    def total_rate(self):
        return lib().ora_total_rate(self._h)

    def select(self, dice):
        s = C.c_int(); m = C.c_int()
        rc = lib().ora_select(self._h, float(dice), C.byref(s), C.byref(m))
        return rc, s.value, m.value

    def transform(self, mode, site):
        lib().ora_transform(self._h, int(mode), int(site))

    def elastic_energy(self):
        return lib().ora_elastic_energy(self._h)

    def set_print(self, printperiod, endstrain):
        lib().ora_set_print(self._h, int(printperiod), float(endstrain))

    def step(self):
        info = OraStepInfo()
        rc = lib().ora_step(self._h, C.byref(info))
        return rc, info

    def save(self, path):
        """Writes the whole state (fields, scalars, RNG, counters, print schedule)."""
        rc = lib().ora_save(self._h, os.fsencode(path))
        if rc:
            raise RuntimeError(f"ora_save({path}) failed: {rc}")

    def load(self, path):
        """Replaces the whole state by an image of ora_save (same parameters)."""
        rc = lib().ora_load(self._h, os.fsencode(path))
        if rc:
            raise RuntimeError(f"ora_load({path}) failed: {rc}")


def set_threads(n):
    """OpenMP threads of the oracle; results do not depend on the count."""
    return lib().ora_set_threads(int(n))


def mt_stream(seed, n):
    r = OraMT()
    L = lib()
    L.ora_mt_seed(C.byref(r), seed)
    return np.array([L.ora_mt_next(C.byref(r)) for _ in range(n)], dtype=np.uint64)


def fft3d(a, sign):
    """Unnormalised complex FFT of a C-ordered array of shape (n2, n1, n0)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    shape = a.shape
    n = list(shape[::-1]) + [1] * (3 - len(shape))
    buf = a.view(np.float64).copy()
    rc = lib().ora_fft3d(buf.ctypes.data_as(C.POINTER(C.c_double)), n[0], n[1], n[2], sign)
    if rc:
        raise RuntimeError("ora_fft3d failed")
    return buf.view(np.complex128).reshape(shape)
