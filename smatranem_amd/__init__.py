"""smatranem_amd -- MI355X-native split-field FDTD time-stepper (drop-in for SMatrAnEM's
FDTD executable, FDTDcalc/FDTDsplitfieldMetal.c).

The compute path is libsmx_fdtd.so (hand-written HIP kernels for gfx950 behind the C
ABI in include/smx_fdtd.h).  This package is a thin ctypes mirror of that ABI; it
fails loudly when the library is missing or no GPU is visible -- there is no CPU
fallback.
"""
from .fdtd import (  # noqa: F401
    F32,
    F64,
    Grid,
    Plan,
    SmxError,
    build_coeffs,
    expn_table,
    lib,
    library_path,
    main_calc,
    remove_sigma,
    set_address,
    set_env,
    set_sigma,
    set_tau,
)
