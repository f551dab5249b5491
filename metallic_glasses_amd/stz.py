"""Python mirror of the C-ABI in include/stz.h (ctypes over libstz.so).

The names follow the reference's routines (STZ/ = "Mesoscale simulation STZ model/"):
stress_solve ~ InhomogeneousStressSolver, rates ~ SofteningFactor + ActEnergy + KMC
rate total, select ~ the KMC search, apply_event ~ VoxelShearTransform, step ~ one
iteration of the main loop of STZ/bmg.c.  Every call runs the HIP kernels of libstz.so;
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libstz.so")

FP64, MIXED = 0, 1
TENSION, SHEAR = 0, 1
SOLVER_FFT, SOLVER_LOCAL = 0, 1
F_EPS0, F_SIGMA, F_Q0, F_DEPS, F_DEPS_LAST, F_PERM, F_TEMP, F_TLAST, F_MINQ, F_TLOCAL, F_MRO = range(11)
ELASTIC, THERMAL, ATHERMAL = 0, 1, 2
ERRORS = {-1: "selection failed", -2: "HIP error", -3: "bad argument", -4: "NaN in stress solve",
          -5: "unsupported (crystal-like MRO)"}


class StzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"stz error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
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


class StepResult(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("site", C.c_int), ("mode", C.c_int), ("n_transformed", C.c_int),
        ("n_negative", C.c_longlong), ("q_chosen", C.c_double), ("total_rate", C.c_double),
        ("dt_kmc", C.c_double), ("totime", C.c_double), ("epsavg", C.c_double * 6),
        ("stress", C.c_double * 6), ("printed", C.c_int), ("soc_elast", C.c_double),
    ]


# the reference's defaults: STZ/input, STZ/bmg.h:22-98, STZ/init.c:60-68
REFERENCE_2D = dict(
    dim=2, n=(512, 512, 1), modes=20, precision=FP64, loading=TENSION, athermal_correction=0,
    softening=1, seed=123, E=69.46, nu=0.365, double_nu=1, q0_mean=2.0, q0_std=0.25,
    gamma0=0.14, temperature=300.0, trial_frequency=1e13, lx=0.7, ly=0.7, lz=1.0,
    strainrate=1e-4, strainincrement=1e-4, perm_soft=13.0, temp_soft=39.0,
    perm_cap=float(-np.log(0.6)), temp_soft_barrier=0.37, athermal_threshold=0.33,
    mro_fraction=0.0, mro_q0=2.0, random_ratio=0.2,
)


def config_3d(n=(64, 64, 64), seed=1234, q0_std=0.2, precision=FP64, **kw):
    """3-D simple-shear workload: mu = 35 GPa, nu = 0.35, M = 20, gamma0 = 0.1, T = 300 K."""
    d = dict(REFERENCE_2D)
    d.update(dim=3, n=tuple(n), E=94.5, nu=0.35, double_nu=0, q0_mean=2.0, q0_std=q0_std,
             gamma0=0.1, seed=seed, loading=SHEAR, athermal_correction=1, precision=precision)
    d.update(kw)
    return d


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m metallic_glasses_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp = C.c_void_p
        dp = P(C.c_double)
        L.stz_create.argtypes = [P(Params), P(vp)]
        L.stz_destroy.argtypes = [vp]
        L.stz_destroy.restype = None
        L.stz_last_error.argtypes = [vp]
        L.stz_last_error.restype = C.c_char_p
        L.stz_init_state.argtypes = [vp]
        L.stz_upload.argtypes = [vp, C.c_int, dp]
        L.stz_download.argtypes = [vp, C.c_int, dp]
        L.stz_set_scalars.argtypes = [vp, dp, dp, C.c_double]
        L.stz_get_scalars.argtypes = [vp, dp, dp, dp]
        L.stz_stress_solve.argtypes = [vp]
        L.stz_rates.argtypes = [vp, P(C.c_longlong), dp]
        L.stz_select.argtypes = [vp, C.c_double, P(C.c_int), P(C.c_int), dp]
        L.stz_apply_event.argtypes = [vp, C.c_int, C.c_int, dp]
        L.stz_athermal_scan.argtypes = [vp, P(C.c_int)]
        L.stz_athermal_list.argtypes = [vp, P(C.c_int), P(C.c_int), C.c_int]
        L.stz_athermal_apply.argtypes = [vp, dp, dp]
        L.stz_add_uniform_stress.argtypes = [vp, C.c_int, C.c_double]
        L.stz_elastic_energy.argtypes = [vp, dp]
        L.stz_set_print.argtypes = [vp, C.c_int, C.c_double]
        L.stz_step.argtypes = [vp, P(StepResult)]
        L.stz_step_kmc.argtypes = [vp, P(StepResult)]
        L.stz_step_print.argtypes = [vp, P(StepResult)]
        L.stz_run_steps.argtypes = [vp, C.c_int, P(StepResult), P(C.c_int)]
        L.stz_rng_seed.argtypes = [vp, C.c_ulong]
        L.stz_rng_uniform.argtypes = [vp, dp]
        L.stz_rng_gaussian.argtypes = [vp, C.c_double, dp]
        L.stz_counters.argtypes = [vp, P(C.c_longlong)]
        L.stz_rng_get_state.argtypes = [vp, P(C.c_uint)]
        L.stz_rng_set_state.argtypes = [vp, P(C.c_uint)]
        L.stz_get_print_state.argtypes = [vp, dp]
        L.stz_set_print_state.argtypes = [vp, C.c_double]
        L.stz_timing_enable.argtypes = [vp, C.c_int]
        L.stz_timing_get.argtypes = [vp, dp, P(C.c_longlong)]
        L.stz_timing_solve_get.argtypes = [vp, dp, P(C.c_longlong)]
        L.stz_run_config.argtypes = [vp, C.c_int, C.c_int, C.c_longlong]
        L.stz_autotune.argtypes = [vp, C.c_int, C.c_int, dp]
        L.stz_init_state_device.argtypes = [vp, C.c_ulonglong, dp, C.c_int, C.c_double, C.c_double, P(C.c_longlong)]
        L.stz_get_tuning.argtypes = [vp, P(C.c_int)]
        L.stz_set_tuning.argtypes = [vp, P(C.c_int)]
        L.stz_tune_candidates.argtypes = [vp, C.c_int, P(C.c_int)]
        L.stz_run_stop_at_print.argtypes = [vp, C.c_int]
        L.stz_ensemble_create.argtypes = [P(vp), C.c_int, P(vp)]
        L.stz_ensemble_config.argtypes = [vp, C.c_int, C.c_int]
        L.stz_ensemble_run.argtypes = [vp, C.c_int, P(P(StepResult)), P(C.c_int), P(C.c_int)]
        L.stz_ensemble_destroy.argtypes = [vp]
        L.stz_ensemble_destroy.restype = None
        L.stz_ensemble_persistent.argtypes = [vp, C.c_int]
        L.stz_run_persistent.argtypes = [vp, C.c_int]
        L.stz_mt_window_probe.argtypes = [C.c_ulong, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                                          dp, dp, P(C.c_longlong), P(C.c_longlong)]
        L.stz_copy_probe.argtypes = [C.c_longlong, C.c_int, dp, dp]
        L.stz_rates_noalloc.argtypes = [P(Params), P(RateFields), C.c_double, vp, C.c_void_p, C.c_void_p]
        L.stz_stress_update_local.argtypes = [vp, C.c_int, P(C.c_int)]
        L.stz_set_solver.argtypes = [vp, C.c_int, C.c_longlong, C.c_int]
        L.stz_checkpoint_save.argtypes = [vp, C.c_char_p]
        L.stz_displacement.argtypes = [vp, dp]
        L.stz_checkpoint_load.argtypes = [vp, C.c_char_p]
        L.stz_set_moduli.argtypes = [vp, dp, dp]
        L.stz_init_state_mro.argtypes = [vp, P(C.c_int), dp]
        L.stz_set_heat_conduct.argtypes = [vp, C.c_int]
        L.stz_heat_conduct.argtypes = [vp, C.c_double]
        L.stz_inhom_iterations.argtypes = [vp]
        _lib = L
    return _lib


def make_params(**kw):
    d = dict(REFERENCE_2D)
    d.update(kw)
    p = Params()
    for k, v in d.items():
        if k == "n":
            for i in range(3):
                p.n[i] = int(v[i]) if i < len(v) else 1
        else:
            setattr(p, k, v)
    if d["dim"] == 2:
        p.n[2] = 1
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Stz:
    """One simulation context on the current HIP device."""

    def __init__(self, init_state=True, device_init=None, mro_map=None, **kw):
        """init_state: draw the reference's MT19937 initial state on the host;
        device_init = dict(seed=..., spheres=[(x, y, z, r), ...], inc_mean=..., inc_std=...):
        draw it on the GPU instead (counter-based, for large grids);
        mro_map = (types[N], thetas[N]): the MRO map of an MROstructure.ms file (InitMRO 1)."""
        self.params = make_params(**kw)
        self._h = C.c_void_p()
        rc = lib().stz_create(C.byref(self.params), C.byref(self._h))
        if rc:
            raise StzError(rc, "stz_create")
        p = self.params
        self.dim = p.dim
        self.dims = (p.n[0], p.n[1], p.n[2])
        self.N = p.n[0] * p.n[1] * p.n[2]
        self.C = 3 if p.dim == 2 else 6
        self.S = 2 if p.dim == 2 else 5
        self.M = p.modes
        self.inclusion_voxels = 0
        if device_init is not None:
            sph = np.ascontiguousarray(np.array(device_init.get("spheres", []), dtype=np.float64).reshape(-1, 4))
            cnt = C.c_longlong()
            self._check(lib().stz_init_state_device(
                self._h, int(device_init.get("seed", p.seed)), _dp(sph) if len(sph) else None, len(sph),
                float(device_init.get("inc_mean", 0.0)), float(device_init.get("inc_std", 0.0)), C.byref(cnt)))
            self.inclusion_voxels = cnt.value
        elif mro_map is not None:
            t = np.ascontiguousarray(mro_map[0], dtype=np.int32).reshape(-1)
            th = np.ascontiguousarray(mro_map[1], dtype=np.float64).reshape(-1)
            if t.size != self.N or th.size != self.N:
                raise ValueError("the MRO map needs one type and one angle per voxel")
            self._check(lib().stz_init_state_mro(self._h, t.ctypes.data_as(C.POINTER(C.c_int)), _dp(th)))
        elif init_state:
            self._check(lib().stz_init_state(self._h))

    def _check(self, rc):
        if rc < 0:
            raise StzError(rc, lib().stz_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().stz_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def size(self, field):
        return {F_EPS0: self.C * self.N, F_SIGMA: self.C * self.N, F_Q0: self.M * self.N,
                F_DEPS: self.M * self.S * self.N, F_DEPS_LAST: self.S * self.N}.get(field, self.N)

    def upload(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size != self.size(field):
            raise ValueError("size mismatch")
        self._check(lib().stz_upload(self._h, field, _dp(a)))

    def download(self, field):
        a = np.empty(self.size(field), dtype=np.float64)
        self._check(lib().stz_download(self._h, field, _dp(a)))
        return a

    def set_scalars(self, epsavg, stress, totime):
        e = np.ascontiguousarray(epsavg, dtype=np.float64)
        s = np.ascontiguousarray(stress, dtype=np.float64)
        self._check(lib().stz_set_scalars(self._h, _dp(e), _dp(s), float(totime)))

    def scalars(self):
        e = np.zeros(6)
        s = np.zeros(6)
        t = C.c_double()
        self._check(lib().stz_get_scalars(self._h, _dp(e), _dp(s), C.byref(t)))
        return e, s, t.value

    def stress_solve(self):
        self._check(lib().stz_stress_solve(self._h))

    def stress_update_local(self, sites):
        """Event-local stress update after transforms at `sites` (see include/stz.h)."""
        a = (C.c_int * max(len(sites), 1))(*[int(x) for x in sites])
        self._check(lib().stz_stress_update_local(self._h, len(sites), a))

    def displacement(self):
        """Print-step displacement field u[dim][N] (see include/stz.h)."""
        u = np.empty(self.dim * self.N, dtype=np.float64)
        self._check(lib().stz_displacement(self._h, _dp(u)))
        return u

    def set_moduli(self, lam, mu):
        """Per-voxel Lambda(x), Mu(x) (arrays of N), or None for uniform moduli (see
        include/stz.h)."""
        if lam is None and mu is None:
            self._check(lib().stz_set_moduli(self._h, None, None))
            return
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        if lam.size != self.N or mu.size != self.N:
            raise ValueError("moduli must have one value per voxel")
        self._check(lib().stz_set_moduli(self._h, _dp(lam), _dp(mu)))

    def set_heat_conduct(self, on=True):
        """HEAT_CONDUCT: diffuse F_TLOCAL after thermal and elastic steps (see include/stz.h)."""
        self._check(lib().stz_set_heat_conduct(self._h, 1 if on else 0))

    def heat_conduct(self, dt):
        """One HeatConducting(dt) of the local temperature field."""
        self._check(lib().stz_heat_conduct(self._h, float(dt)))

    def lame(self):
        """(lambda, mu) of the reference medium, derived from the parameters as the library
        derives them (STZ/init.c:26-44: nu doubled with double_nu, plane strain in 2-D)."""
        p = self.params
        nu = p.nu * 2 if p.double_nu else p.nu
        mu = p.E / (2 + 2 * nu)
        lam = p.E * nu / (1 - nu * nu) if p.dim == 2 else p.E * nu / ((1 + nu) * (1 - 2 * nu))
        return lam, mu

    def inhom_iterations(self):
        """Homogeneous solves the last stress solve took (1 with uniform moduli)."""
        return int(lib().stz_inhom_iterations(self._h))

    def checkpoint_save(self, path):
        """Writes everything the next step reads (see include/stz.h)."""
        self._check(lib().stz_checkpoint_save(self._h, os.fsencode(path)))

    def checkpoint_load(self, path):
        self._check(lib().stz_checkpoint_load(self._h, os.fsencode(path)))

    def set_solver(self, mode, resync=1024, local_max=16):
        """Stress update of the resident step: SOLVER_FFT or SOLVER_LOCAL."""
        self._check(lib().stz_set_solver(self._h, int(mode), int(resync), int(local_max)))

    def rates(self):
        n = C.c_longlong()
        r = C.c_double()
        self._check(lib().stz_rates(self._h, C.byref(n), C.byref(r)))
        return n.value, r.value

    def select(self, dice):
        s = C.c_int()
        m = C.c_int()
        q = C.c_double()
        rc = lib().stz_select(self._h, float(dice), C.byref(s), C.byref(m), C.byref(q))
        return rc, s.value, m.value, q.value

    def apply_event(self, site, mode, deps_new):
        a = np.ascontiguousarray(deps_new, dtype=np.float64)
        self._check(lib().stz_apply_event(self._h, int(site), int(mode), _dp(a)))

    def athermal_scan(self):
        n = C.c_int()
        self._check(lib().stz_athermal_scan(self._h, C.byref(n)))
        return n.value

    def athermal_list(self, n):
        s = np.zeros(max(n, 1), dtype=np.int32)
        m = np.zeros(max(n, 1), dtype=np.int32)
        self._check(lib().stz_athermal_list(self._h, s.ctypes.data_as(C.POINTER(C.c_int)),
                                            m.ctypes.data_as(C.POINTER(C.c_int)), int(n)))
        return s[:n], m[:n]

    def athermal_apply(self, deps_new):
        a = np.ascontiguousarray(deps_new, dtype=np.float64)
        soc = C.c_double()
        self._check(lib().stz_athermal_apply(self._h, _dp(a) if a.size else None, C.byref(soc)))
        return soc.value

    def add_uniform_stress(self, comp, delta):
        self._check(lib().stz_add_uniform_stress(self._h, int(comp), float(delta)))

    def elastic_energy(self):
        e = C.c_double()
        self._check(lib().stz_elastic_energy(self._h, C.byref(e)))
        return e.value

    def set_print(self, printperiod, endstrain):
        self._check(lib().stz_set_print(self._h, int(printperiod), float(endstrain)))

    def step(self):
        r = StepResult()
        rc = self._check(lib().stz_step(self._h, C.byref(r)))
        return rc, r

    def run_steps(self, n):
        """Runs n device-resident steps; returns (rc, results[:done])."""
        arr = (StepResult * max(n, 1))()
        done = C.c_int()
        rc = self._check(lib().stz_run_steps(self._h, int(n), arr, C.byref(done)))
        return rc, arr[:done.value]

    def run_config(self, batch=0, use_graph=-1, window_words=0):
        self._check(lib().stz_run_config(self._h, int(batch), int(use_graph), int(window_words)))

    def persistent(self, on=True):
        """Resident steps as one persistent launch per step (stz_run_persistent)."""
        self._check(lib().stz_run_persistent(self._h, int(bool(on))))

    def autotune(self, reps=10, force=False):
        """Times the solve kernels' tile variants and keeps the fastest; returns
        (choices, mean ms of the chosen variants)."""
        ms = np.zeros(4)
        self._check(lib().stz_autotune(self._h, int(reps), 1 if force else 0, _dp(ms)))
        return self.tuning(), ms

    def tuning(self):
        o = (C.c_int * 4)()
        self._check(lib().stz_get_tuning(self._h, o))
        return list(o)

    def tune_candidates(self, slot):
        o = (C.c_int * 16)()
        n = self._check(lib().stz_tune_candidates(self._h, int(slot), o))
        return list(o)[:n]

    def set_tuning(self, t):
        a = (C.c_int * 4)(*[int(x) for x in t])
        self._check(lib().stz_set_tuning(self._h, a))

    def stop_at_print(self, on=True):
        self._check(lib().stz_run_stop_at_print(self._h, 1 if on else 0))

    def step_kmc(self):
        r = StepResult()
        rc = self._check(lib().stz_step_kmc(self._h, C.byref(r)))
        return rc, r

    def step_print(self, r):
        return self._check(lib().stz_step_print(self._h, C.byref(r)))

    def counters(self):
        o = np.zeros(6, dtype=np.int64)
        self._check(lib().stz_counters(self._h, o.ctypes.data_as(C.POINTER(C.c_longlong))))
        return o

    def rng_state(self):
        w = np.zeros(625, dtype=np.uint32)
        self._check(lib().stz_rng_get_state(self._h, w.ctypes.data_as(C.POINTER(C.c_uint))))
        return w

    def set_rng_state(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        self._check(lib().stz_rng_set_state(self._h, w.ctypes.data_as(C.POINTER(C.c_uint))))

    def print_state(self):
        o = np.zeros(3)
        self._check(lib().stz_get_print_state(self._h, _dp(o)))
        return o

    def set_print_state(self, strainhist_last):
        self._check(lib().stz_set_print_state(self._h, float(strainhist_last)))

    def timing(self, on=True):
        self._check(lib().stz_timing_enable(self._h, 1 if on else 0))

    def timing_get(self):
        ms = np.zeros(8)
        n = np.zeros(8, dtype=np.int64)
        self._check(lib().stz_timing_get(self._h, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_longlong))))
        ms2 = np.zeros(8)
        n2 = np.zeros(8, dtype=np.int64)
        self._check(lib().stz_timing_solve_get(self._h, _dp(ms2), n2.ctypes.data_as(C.POINTER(C.c_longlong))))
        return ms, n, ms2, n2

    def rng_seed(self, seed):
        self._check(lib().stz_rng_seed(self._h, int(seed)))

    def rng_uniform(self):
        u = C.c_double()
        self._check(lib().stz_rng_uniform(self._h, C.byref(u)))
        return u.value

    def rng_gaussian(self, sigma=1.0):
        g = C.c_double()
        self._check(lib().stz_rng_gaussian(self._h, float(sigma), C.byref(g)))
        return g.value


class Ensemble:
    """Replicas advanced together by the batched resident step (stz_ensemble_*)."""

    def __init__(self, members, batch=0, use_graph=-1, persistent=False):
        self.members = list(members)
        hs = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = C.c_void_p()
        rc = lib().stz_ensemble_create(hs, len(self.members), C.byref(self._h))
        if rc:
            raise StzError(rc, "stz_ensemble_create")
        if batch or use_graph >= 0:
            rc = lib().stz_ensemble_config(self._h, int(batch), int(use_graph))
            if rc:
                raise StzError(rc, "stz_ensemble_config")
        if persistent:
            self.persistent(True)

    def persistent(self, on=True):
        """One persistent launch per step for all replicas (stz_ensemble_persistent)."""
        rc = lib().stz_ensemble_persistent(self._h, int(bool(on)))
        if rc:
            raise StzError(rc, lib().stz_last_error(self.members[0]._h).decode())

    def run(self, n, keep=True):
        """Advances every replica by up to n steps; returns (rcs, [results per replica])."""
        k = len(self.members)
        bufs = [(StepResult * max(n, 1))() for _ in range(k)] if keep else None
        ptrs = (C.POINTER(StepResult) * k)(*[C.cast(b, C.POINTER(StepResult)) for b in bufs]) if keep else None
        done = (C.c_int * k)()
        rcs = (C.c_int * k)()
        rc = lib().stz_ensemble_run(self._h, int(n), ptrs, done, rcs)
        if rc:
            raise StzError(rc, lib().stz_last_error(self.members[0]._h).decode())
        out = [bufs[i][:done[i]] for i in range(k)] if keep else [done[i] for i in range(k)]
        return list(rcs), out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().stz_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


def mt_window_probe(seed, skip, window_words, offset, with_uniform, ngauss):
    """Lookup-window draws vs the same draws in order (host only)."""
    n = ngauss + (1 if with_uniform else 0)
    w = np.zeros(max(n, 1))
    s = np.zeros(max(n, 1))
    we = C.c_longlong()
    se = C.c_longlong()
    rc = lib().stz_mt_window_probe(int(seed), int(skip), int(window_words), int(offset), int(with_uniform),
                                   int(ngauss), _dp(w), _dp(s), C.byref(we), C.byref(se))
    if rc:
        raise StzError(rc, "stz_mt_window_probe")
    return w[:n], s[:n], we.value, se.value


class RateFields(C.Structure):
    """Device pointers of stz_rates_noalloc's fields (include/stz.h stz_rate_fields)."""
    _fields_ = [(k, C.c_void_p) for k in ("sigma", "q0", "deps", "deps_last", "perm", "temp", "tlast")]


def rates_noalloc(kw, ptrs, totime, total_rate_ptr, n_negative_ptr, stream=None):
    """stz_rates_noalloc: R and the Q<0 count of FP64 fields in device memory (ptrs: a dict
    of device addresses by RateFields name), written to device memory at total_rate_ptr
    (double) and n_negative_ptr (int64) when the kernels run on `stream` (None: the null
    stream).  Returns the C return code."""
    p = make_params(**kw)
    f = RateFields(**{k: C.c_void_p(int(v)) for k, v in ptrs.items()})
    return lib().stz_rates_noalloc(C.byref(p), C.byref(f), float(totime), stream,
                                   C.c_void_p(int(total_rate_ptr)), C.c_void_p(int(n_negative_ptr)))


def copy_probe(nbytes, reps=10):
    """Plain device copy of nbytes: (ms per launch, GB/s counting read + write)."""
    ms = C.c_double()
    gbs = C.c_double()
    rc = lib().stz_copy_probe(int(nbytes), int(reps), C.byref(ms), C.byref(gbs))
    if rc:
        raise StzError(rc, "stz_copy_probe")
    return ms.value, gbs.value


def soft_inclusions(n, fraction=0.05, radius=16.0, seed=1234):
    """Sphere centres for a soft-inclusion region: count = fraction of the grid volume
    over the sphere volume, centres uniform on the periodic grid from MT19937(seed)."""
    vol = float(n[0]) * n[1] * n[2]
    k = max(1, int(round(fraction * vol / (4.0 / 3.0 * np.pi * radius ** 3))))
    rs = np.random.RandomState(seed)
    c = rs.uniform(0.0, 1.0, size=(k, 3)) * np.array(n[:3], dtype=np.float64)
    return [(float(x), float(y), float(z), float(radius)) for x, y, z in c]
