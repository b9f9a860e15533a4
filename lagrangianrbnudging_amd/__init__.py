"""MI355X-native Rayleigh–Bénard lattice-Boltzmann step with Lagrangian particles and nudging.

The computation lives in librbn.so (HIP kernels for gfx950 behind the C-ABI of include/rbn.h);
this package holds its sources (csrc/), the build recipe (build.py), the C host driver
(driver/) and a ctypes binding (rbn.py).
"""
from .rbn import (  # noqa: F401
    RBN_FP32,
    RBN_FP64,
    RBN_VARIANT_NUDGING,
    RBN_VARIANT_TRAJECTORIES,
    Params,
    RbnError,
    Solver,
    default_params,
    device_count,
    load,
    params_from_ra,
    parse_params,
    parse_params_text,
    rayleigh,
)
