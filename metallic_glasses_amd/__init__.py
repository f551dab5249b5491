"""MI355X-native STZ kinetic Monte Carlo step (hand-written HIP for gfx950 behind a C-ABI).

The compute path lives in libstz.so (built from csrc/ by metallic_glasses_amd.build);
stz.py is the Python mirror of include/stz.h.
"""
from .stz import (ATHERMAL, ELASTIC, F_DEPS, F_DEPS_LAST, F_EPS0, F_MINQ, F_PERM, F_Q0, F_SIGMA,  # noqa: F401
                  F_TEMP, F_TLAST, F_TLOCAL, F_MRO, FP64, MIXED, REFERENCE_2D, SHEAR, SOLVER_FFT, SOLVER_LOCAL, TENSION, THERMAL, StepResult, Stz, StzError,
                  Ensemble, RateFields, config_3d, copy_probe, lib, make_params, mt_window_probe, rates_noalloc,
                  soft_inclusions)
