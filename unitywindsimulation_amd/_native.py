"""ctypes binding of the C ABI declared in ``include/uws.h``.

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libuws_hip.so"

# ---- constants mirrored from uws.h ----
UWS_OK = 0
UWS_ERR_INVALID_ARG, UWS_ERR_CONFIG, UWS_ERR_ALLOC, UWS_ERR_DEVICE, UWS_ERR_UNSUPPORTED, UWS_ERR_NONFINITE = -1, -2, -3, -4, -5, -6
UWS_ERR_TOO_LARGE = -7
BG_STAGING_MAX_BYTES = 32 << 20
VEL_FIXED_VALUE, VEL_ZERO_GRAD, VEL_SYMMETRY = 0, 1, 2
PRES_FIXED_VALUE, PRES_ZERO_GRAD, PRES_SYMMETRY = 0, 1, 2
SIMPLE, PISO = 0, 1
TURB_NONE, TURB_SMAGORINSKY = 0, 1
INFLOW_CONSTANT, INFLOW_LOG, INFLOW_POWER_LAW = 0, 1, 2
CELL_FLUID, CELL_SOLID = 0, 1
WALL_SMOOTH, WALL_ROUGH = 0, 1
X0, XN, Y0, YN, Z0, ZN = range(6)
GRID_COLLOCATED, GRID_COLL_NONUNIFORM, GRID_STAGGERED = 0, 1, 2

FIELD_VEL, FIELD_PRES, FIELD_VEL_WORK, FIELD_VEL_CORR, FIELD_PRES_CORR = 0, 1, 2, 3, 4
FIELD_EDDY_VIS, FIELD_WALL_VIS, FIELD_B, FIELD_D, FIELD_DP, FIELD_BP, FIELD_AODU = 5, 6, 7, 8, 9, 10, 11
FIELD_FACE_VEL = 12
FIELD_NCOMP = {0: 3, 1: 1, 2: 3, 3: 3, 4: 1, 5: 1, 6: 3, 7: 3, 8: 3, 9: 1, 10: 1, 11: 3, 12: 3}

KERNEL_VEL_SWEEP, KERNEL_PRES_SWEEP, KERNEL_PRES_SWEEP2, KERNEL_VEL_SWEEP2, KERNEL_PRES_SWEEP3, KERNEL_PRES_SWEEP4 = \
    0, 1, 2, 3, 4, 5
PREC_FP32, PREC_FP64, PREC_FP32_FP16, PREC_FP32_BF16, PREC_CSTEP = 0, 1, 2, 3, 4
CSTEP_VEL, CSTEP_PRES, CSTEP_BND, CSTEP_CS = 0, 1, 2, 3
GRAD_VEL0, GRAD_PRES0, GRAD_BND, GRAD_CS = 0, 1, 2, 3
NORM_VEL_RESIDUAL, NORM_PRES_RESIDUAL, NORM_DIVERGENCE, NORM_NONFINITE, NORM_REFINE0, NORM_COUNT = 0, 1, 2, 3, 4, 64

(STAGE_WALL_FUNC, STAGE_EDDY_VIS, STAGE_LES_TERM, STAGE_VEL_PRECOMP, STAGE_VEL_SWEEPS, STAGE_PRES_PRECOMP,
 STAGE_PRES_SWEEPS, STAGE_APPLY_PRES, STAGE_VEL_CORRECT, STAGE_PISO, STAGE_FINISH) = range(11)


class uws_background(ctypes.Structure):
    """Mirror of ``struct uws_background`` (include/uws.h)."""
    _fields_ = [("grid_res", c_int32 * 3), ("dx", c_float), ("phys_field_pos", c_float * 3)]


class uws_config(ctypes.Structure):
    """Mirror of ``struct uws_config`` (include/uws.h)."""

    _fields_ = [
        ("grid_res", c_int32 * 3),
        ("dx", c_float),
        ("dt", c_float),
        ("mu", c_float),
        ("den", c_float),
        ("external_force", c_float * 3),
        ("vel_bnd", c_int32 * 6),
        ("pres_bnd", c_int32 * 6),
        ("fvm_solver_type", c_int32),
        ("piso_num_correctors", c_int32),
        ("turbulence_model", c_int32),
        ("smagorinsky_constant", c_float),
        ("vel_max_num_iter", c_int32),
        ("pres_max_num_iter", c_int32),
        ("cal_residual", c_int32),
        ("vel_residual_check_interval", c_int32),
        ("vel_tolerance", c_float),
        ("pres_residual_check_interval", c_int32),
        ("pres_tolerance", c_float),
        ("wall_slots", c_int32),
        ("wall_model", c_int32),
        ("wall_z0", c_float),
        ("precision", c_int32),
        ("refine_rounds", c_int32),
        ("refine_sweeps", c_int32),
        ("compute_norms", c_int32),
        ("grid_type", c_int32),
        ("reserved", c_int32 * 3),
    ]


class UwsError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} ({status})")


# every exported symbol with (restype, argtypes); kept in the order of uws.h
SIGNATURES = {
    "uws_config_default": (c_int, [POINTER(uws_config)]),
    "uws_strerror": (c_char_p, [c_int]),
    "uws_abi_version": (c_int, []),
    "uws_create": (c_int, [POINTER(uws_config), POINTER(c_void_p)]),
    "uws_destroy": (c_int, [c_void_p]),
    "uws_change_phys_domain_size": (c_int, [c_void_p, POINTER(uws_config)]),
    "uws_get_config": (c_int, [c_void_p, POINTER(uws_config)]),
    "uws_num_cells": (c_int, [c_void_p, POINTER(c_int64)]),
    "uws_init_flags": (c_int, [c_void_p, c_void_p, c_int64]),
    "uws_init_box": (c_int, [c_void_p]),
    "uws_set_face_positions": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "uws_get_face_positions": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "uws_stretched_face_positions": (c_int, [c_int32, c_float, c_int32, c_int32, c_float, c_void_p]),
    "uws_init_vel_pres_fields": (c_int, [c_void_p, c_int32, c_float]),
    "uws_set_fixed_value_vel_bnd_cond": (c_int, [c_void_p, c_int32, c_float, c_void_p, c_float, c_float]),
    "uws_set_bnd_field": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_get_bnd_field": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_set_velocity": (c_int, [c_void_p, c_void_p, c_int32, c_int64]),
    "uws_get_velocity": (c_int, [c_void_p, c_void_p, c_int32, c_int64]),
    "uws_set_pressure": (c_int, [c_void_p, c_void_p, c_int64]),
    "uws_get_pressure": (c_int, [c_void_p, c_void_p, c_int64]),
    "uws_get_flags": (c_int, [c_void_p, c_void_p, c_int64]),
    "uws_init_noise": (c_int, [c_void_p, c_uint64, c_float]),
    "uws_step": (c_int, [c_void_p, c_int32]),
    "uws_synchronize": (c_int, [c_void_p]),
    "uws_get_time": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_double)]),
    "uws_field_ncomp": (c_int, [c_int32]),
    "uws_get_field": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_set_field": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_run_stage": (c_int, [c_void_p, c_int32]),
    "uws_get_field_f64": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_get_norms": (c_int, [c_void_p, c_void_p, c_int32]),
    "uws_get_sweep_counts": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "uws_get_launch_plan": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "uws_stream_bandwidth": (c_int, [c_int64, c_int32, c_int32, POINTER(c_double)]),
    "uws_time_kernel": (c_int, [c_void_p, c_int32, c_int32, POINTER(c_double)]),
    "uws_batch_create": (c_int, [c_void_p, c_int32, POINTER(c_void_p)]),
    "uws_batch_destroy": (c_int, [c_void_p]),
    "uws_batch_num_cases": (c_int, [c_void_p, POINTER(c_int32)]),
    "uws_batch_init_flags": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_batch_init_noise": (c_int, [c_void_p, c_int32, c_uint64, c_float]),
    "uws_batch_set_fixed_value_vel_bnd_cond": (c_int, [c_void_p, c_int32, c_int32, c_float, c_void_p, c_float,
                                                       c_float]),
    "uws_batch_set_velocity": (c_int, [c_void_p, c_int32, c_void_p, c_int32, c_int64]),
    "uws_batch_get_velocity": (c_int, [c_void_p, c_int32, c_void_p, c_int32, c_int64]),
    "uws_batch_get_pressure": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_batch_get_norms": (c_int, [c_void_p, c_int32, c_void_p, c_int32]),
    "uws_batch_step": (c_int, [c_void_p, c_int32]),
    "uws_batch_synchronize": (c_int, [c_void_p]),
    "uws_batch_set_launch": (c_int, [c_void_p, c_int32, c_int32, c_int64]),
    "uws_batch_num_groups": (c_int, [c_void_p, POINTER(c_int32)]),
    "uws_batch_autotune": (c_int, [c_void_p, c_int32, c_void_p, POINTER(c_double)]),
    "uws_init_vel_pres_fields_from_background": (c_int, [c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_int32,
                                                         c_void_p]),
    "uws_init_vel_pres_fields_from_background_noalloc": (c_int, [c_void_p, c_void_p, c_float, c_void_p, c_void_p,
                                                                 c_int32, c_void_p]),
    "uws_mean_reset": (c_int, [c_void_p]),
    "uws_mean_accumulate": (c_int, [c_void_p]),
    "uws_get_mean_velocity": (c_int, [c_void_p, c_void_p, c_int64, POINTER(c_int64)]),
    "uws_set_smagorinsky_constant": (c_int, [c_void_p, c_double]),
    "uws_set_bnd_field_f64": (c_int, [c_void_p, c_int32, c_void_p, c_int64]),
    "uws_cstep_set_imag": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_int64]),
    "uws_cstep_get_imag": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_int64]),
    "uws_adjoint_create": (c_int, [c_void_p, c_int32, c_int32, POINTER(c_void_p)]),
    "uws_adjoint_destroy": (c_int, [c_void_p]),
    "uws_adjoint_forward": (c_int, [c_void_p]),
    "uws_adjoint_backward": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64]),
    "uws_adjoint_get_grad": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_int64]),
    "uws_adjoint_memory": (c_int, [c_void_p, POINTER(c_int64)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load ``libuws_hip.so`` (built by ``unitywindsimulation_amd.build``)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64 /
        # libhsa-runtime64 with the same sonames as /opt/rocm's.  Loaded first,
        # torch's copies also serve this library; loaded after it, torch would
        # bring a second runtime into the process and device calls of this
        # library then fail.  So torch, where installed, is imported first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = Path(os.environ.get("UWS_LIB", LIB_PATH))
        if not path.exists():
            raise ImportError(
                f"HIP library {path} is missing: build it with "
                "`python -m unitywindsimulation_amd.build` (there is no CPU fallback)"
            )
        dll = ctypes.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(dll, name)
            fn.restype = res
            fn.argtypes = args
        _lib = dll
    return _lib


def strerror(status: int) -> str:
    try:
        return lib().uws_strerror(status).decode()
    except ImportError:
        return "library not loaded"


def check(status: int, what: str) -> None:
    if status != UWS_OK:
        raise UwsError(status, what)


def ptr(a: np.ndarray) -> c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_void_p)
