"""ctypes binding of ``libpmpp_amd.so`` (the C ABI declared in ``include/pmpp_amd.h``).

This mirrors the reference's native-library loader (``src/pmpp/distributed/cuda.py:136-230``:
locate the shared object, load it with ``ctypes.CDLL``, check an ABI version, expose
typed entry points).  The XLA-FFI registration the reference adds on top of this
(``jax.ffi.register_ffi_target``) needs jaxlib, which this package does not depend on.

There is no fallback: if the library is missing, every operator raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

ABI_VERSION = 9

PM_SUCCESS = 0
PM_ERROR_INVALID_ARGUMENT = 1
PM_ERROR_UNSUPPORTED = 2
PM_ERROR_WORKSPACE = 3
PM_ERROR_HIP = 4

# every symbol the header declares, in header order
EXPORTED_SYMBOLS = (
    "pm_last_error", "pm_abi_version", "pm_init",
    "pm_timing_enable", "pm_timing_mark", "pm_timing_read",
    "pm_cic_scatter", "pm_cic_gather", "pm_cic_gather_bwd", "pm_cic_scatter_bwd",
    "pm_sort_cells_workspace", "pm_sort_cells", "pm_bucket_cells_workspace", "pm_bucket_cells",
    "pm_bucket_cells_noalloc",
    "pm_fft_workspace", "pm_rfft3d", "pm_irfft3d",
    "pm_spectral_workspace", "pm_spectral_force_multiply",
    "pm_gravity_workspace", "pm_gravity_vjp_workspace", "pm_gravity", "pm_gravity_vjp",
    "pm_gravity_keep", "pm_gravity_vjp_resume",
    "pm_gravity_density", "pm_gravity_density_offset", "pm_gravity_corrected", "pm_gravity_corrected_resume", "pm_gravity_corrected_vjp",
    "pm_gravity_density_cot", "pm_potential_workspace", "pm_gravity_potential",
    "pm_kick", "pm_drift", "pm_nbody_adj_workspace", "pm_kick_adj", "pm_drift_adj",
    "pm_power_spectrum", "pm_power_spectrum_bwd",
    "pm_conv3d_fwd", "pm_conv3d_bwd_data", "pm_conv3d_bwd_data_out1", "pm_conv3d_bwd_weight_workspace",
    "pm_conv3d_bwd_weight",
    "pm_conv3d_stem", "pm_conv3d_out_residual", "pm_conv3d_stem_hidden",
    "pm_correction_workspace", "pm_correction_features", "pm_correction_residual", "pm_correction_residual_bwd",
    "pm_correction_input_bwd", "pm_force_loss", "pm_adamax_step", "pm_to_bf16",
    "pm_ic_linear_modes", "pm_ic_spectral_op", "pm_ic_product_acc", "pm_ic_lpt_accumulate",
)


class CicGeom(ctypes.Structure):
    _fields_ = [
        ("local_shape", ctypes.c_int32 * 3),
        ("global_shape", ctypes.c_int32 * 3),
        ("offset", ctypes.c_float * 3),
        ("particle_cell", ctypes.c_float),
        ("mesh_cell", ctypes.c_float),
    ]


class GravityDesc(ctypes.Structure):
    _fields_ = [
        ("mesh_shape", ctypes.c_int32 * 3),
        ("cell_size", ctypes.c_float),
        ("ptcl_spacing", ctypes.c_float),
        ("mask_active", ctypes.c_int32),
        ("omega_m", ctypes.c_float),
        ("particle_value", ctypes.c_float),
        ("brick", ctypes.c_int32 * 3),
    ]


class PMError(RuntimeError):
    """A status code other than PM_SUCCESS from the HIP library."""


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_f32 = ctypes.c_float
_sz_p = ctypes.POINTER(ctypes.c_size_t)
_i32x3 = ctypes.POINTER(ctypes.c_int32)

_SIGNATURES = {
    "pm_last_error": ([], ctypes.c_char_p),
    "pm_abi_version": ([], ctypes.c_int),
    "pm_init": ([], ctypes.c_int),
    "pm_timing_enable": ([_i32], ctypes.c_int),
    "pm_timing_mark": ([ctypes.c_char_p, _vp], ctypes.c_int),
    "pm_timing_read": ([ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_char_p), _i32,
                        ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "pm_cic_scatter": ([_vp, _vp, _vp, _vp, _f32, _i64, _i32, ctypes.POINTER(CicGeom), _vp, _vp], ctypes.c_int),
    "pm_cic_gather": ([_vp, _vp, _vp, _i64, _i32, ctypes.POINTER(CicGeom), _vp, _vp, _vp], ctypes.c_int),
    "pm_cic_gather_bwd": ([_vp, _vp, _vp, _i64, _i32, ctypes.POINTER(CicGeom), _vp, _vp, _vp, _vp, _vp],
                          ctypes.c_int),
    "pm_cic_scatter_bwd": ([_vp, _vp, _vp, _vp, _f32, _i64, _i32, ctypes.POINTER(CicGeom), _vp, _vp, _vp, _vp],
                           ctypes.c_int),
    "pm_sort_cells_workspace": ([_i64, _i32x3, _i32x3, _sz_p], ctypes.c_int),
    "pm_sort_cells": ([_vp, _vp, _vp, _i64, _i32x3, _f32, _i32x3, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
                      ctypes.c_int),
    "pm_bucket_cells_workspace": ([_i64, _i32x3, _i32x3, _sz_p], ctypes.c_int),
    "pm_bucket_cells": ([_vp, _vp, _vp, _i64, _i32x3, _f32, _i32x3, _i32x3, _vp, _vp, _vp, _vp, _vp, _vp,
                         ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_bucket_cells_noalloc": ([_vp, _vp, _vp, _i64, _i32x3, _f32, _i32x3, _i32x3, _vp, _vp, _vp, _vp, _vp, _vp],
                                ctypes.c_int),
    "pm_fft_workspace": ([_i32x3, _sz_p], ctypes.c_int),
    "pm_rfft3d": ([_vp, _vp, _i32x3, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_irfft3d": ([_vp, _vp, _i32x3, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_spectral_workspace": ([ctypes.POINTER(GravityDesc), _sz_p], ctypes.c_int),
    "pm_spectral_force_multiply": ([_vp, _vp, ctypes.POINTER(GravityDesc), _i32, _i32, _vp, ctypes.c_size_t, _vp],
                                   ctypes.c_int),
    "pm_gravity_workspace": ([ctypes.POINTER(GravityDesc), _i64, _sz_p], ctypes.c_int),
    "pm_gravity_density_offset": ([ctypes.POINTER(GravityDesc), _i64, _sz_p], ctypes.c_int),
    "pm_gravity_vjp_workspace": ([ctypes.POINTER(GravityDesc), _i64, _sz_p], ctypes.c_int),
    "pm_gravity": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, ctypes.c_size_t, _vp],
                   ctypes.c_int),
    "pm_gravity_vjp": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
                       ctypes.c_int),
    "pm_gravity_density": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, ctypes.c_size_t, _vp],
                           ctypes.c_int),
    "pm_gravity_keep": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, ctypes.c_size_t, _vp],
                        ctypes.c_int),
    "pm_gravity_vjp_resume": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, _vp,
                               ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_gravity_corrected": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, ctypes.c_size_t,
                              _vp], ctypes.c_int),
    "pm_gravity_corrected_resume": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, ctypes.c_size_t,
                              _vp], ctypes.c_int),
    "pm_gravity_corrected_vjp": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, _vp, _vp, _vp,
                                  ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_gravity_density_cot": ([_vp, _vp, _vp, _i64, ctypes.POINTER(GravityDesc), _vp, _vp, _vp, ctypes.c_size_t,
                                _vp], ctypes.c_int),
    "pm_potential_workspace": ([ctypes.POINTER(GravityDesc), _sz_p], ctypes.c_int),
    "pm_gravity_potential": ([_vp, _f32, _f32, _vp, ctypes.POINTER(GravityDesc), _vp, _vp, ctypes.c_size_t, _vp],
                             ctypes.c_int),
    "pm_kick": ([_vp, _vp, _vp, _i64, _f32, _vp], ctypes.c_int),
    "pm_drift": ([_vp, _vp, _vp, _i64, _f32, _vp], ctypes.c_int),
    "pm_nbody_adj_workspace": ([_sz_p], ctypes.c_int),
    "pm_kick_adj": ([_vp, _vp, _vp, _vp, _i64, _f32, _vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_drift_adj": ([_vp, _vp, _vp, _vp, _i64, _f32, _vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_power_spectrum": ([_vp, _vp, _i32x3, _i32, _i32, _i32, _vp, _vp], ctypes.c_int),
    "pm_power_spectrum_bwd": ([_vp, _i32x3, _i32, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "pm_conv3d_fwd": ([_vp, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _i32x3, _vp], ctypes.c_int),
    "pm_conv3d_bwd_data": ([_vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32x3, _vp], ctypes.c_int),
    "pm_conv3d_bwd_data_out1": ([_vp, _vp, _vp, _vp, _i32, _i32x3, _vp], ctypes.c_int),
    "pm_conv3d_bwd_weight_workspace": ([_i32x3, _i32, _sz_p], ctypes.c_int),
    "pm_conv3d_bwd_weight": ([_vp, _vp, _i32, _i32x3, _vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_conv3d_stem": ([_vp, _vp, _i32x3, _f32, _f32, _f32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "pm_conv3d_out_residual": ([_vp, _vp, _vp, _vp, _i32x3, _f32, _f32, _vp, _vp, _vp], ctypes.c_int),
    "pm_conv3d_stem_hidden": ([_vp, _vp, _i32x3, _f32, _f32, _f32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                              ctypes.c_int),
    "pm_correction_workspace": ([_sz_p], ctypes.c_int),
    "pm_correction_features": ([_vp, _vp, _i64, _f32, _f32, _f32, _f32, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
                               ctypes.c_int),
    "pm_correction_residual": ([_vp, _vp, _i64, _f32, _f32, _vp, _vp, _vp], ctypes.c_int),
    "pm_correction_residual_bwd": ([_vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _vp, _vp, _i32, _vp, _vp,
                                    ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_correction_input_bwd": ([_vp, _vp, _vp, _vp, _vp, _i64, _f32, _f32, _vp, _vp, _vp, _vp, _vp, _vp,
                                 ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_force_loss": ([_vp, _vp, _i64, _vp, _vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "pm_adamax_step": ([_vp, _vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _f32, _vp, _vp, ctypes.c_size_t,
                        _vp], ctypes.c_int),
    "pm_to_bf16": ([_vp, _i64, _vp, _vp], ctypes.c_int),
    "pm_ic_linear_modes": ([_vp, _i32x3, _vp, _vp, _vp, _i32, _f32, _f32, _f32, _f32, _vp], ctypes.c_int),
    "pm_ic_spectral_op": ([_vp, _vp, _i32x3, _vp, _i32, _i32, _i32, _f32, _vp, _vp, _vp], ctypes.c_int),
    "pm_ic_product_acc": ([_vp, _vp, _vp, _i64, _f32, _i32, _vp], ctypes.c_int),
    "pm_ic_lpt_accumulate": ([_vp, _vp, _vp, _i64, _i32, _f32, _f32, _vp], ctypes.c_int),
}

_LIB = None


def library_path() -> Path:
    env = os.environ.get("PMPP_AMD_LIBRARY")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent / "libpmpp_amd.so"


def load_library():
    """Load and type the HIP library once; raise if it is missing or the ABI differs."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise PMError(
            f"pmpp_amd HIP library not found at {path}; build it with `python -m pmpp_amd.build` "
            "(there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (argtypes, restype) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    got = lib.pm_abi_version()
    if got != ABI_VERSION:
        raise PMError(f"pmpp_amd ABI version {got} != expected {ABI_VERSION}")
    _LIB = lib
    return lib


_EXC = {
    PM_ERROR_INVALID_ARGUMENT: ValueError,
    PM_ERROR_UNSUPPORTED: NotImplementedError,
    PM_ERROR_WORKSPACE: PMError,
    PM_ERROR_HIP: PMError,
}


def check(status: int, what: str) -> None:
    if status == PM_SUCCESS:
        return
    msg = load_library().pm_last_error().decode(errors="replace")
    raise _EXC.get(status, PMError)(f"{what}: {msg} (status {status})")


def call(name: str, *args) -> None:
    lib = load_library()
    check(getattr(lib, name)(*args), name)


def int3(values) -> ctypes.Array:
    return (ctypes.c_int32 * 3)(*[int(v) for v in values])


def size_query(name: str, *args) -> int:
    out = ctypes.c_size_t(0)
    call(name, *args, ctypes.byref(out))
    return int(out.value)
