"""ctypes binding of libcfft_hip.so (the C ABI declared in include/cfft_hip.h).

The library is loaded from the package directory (built in-tree by
``crystallographicfft.jl_amd/build.py``).  There is no fallback: when the
library is missing every entry point raises ``CFFTLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcfft_hip.so")

CFFT_OK = 0
CFFT_ERR_INVALID = 1
CFFT_ERR_NOT_COMMENSURATE = 2
CFFT_ERR_NOT_DIVISIBLE = 3
CFFT_ERR_UNREACHABLE = 4
CFFT_ERR_FILL_MAP = 5
CFFT_ERR_CENTERING = 6
CFFT_ERR_ODD_M = 7
CFFT_ERR_SIZE = 8
CFFT_ERR_UNKNOWN_METHOD = 9
CFFT_ERR_BOUNDS = 10
CFFT_ERR_UNSUPPORTED = 11
CFFT_ERR_HIP = 12

PATH_AUTO, PATH_M2, PATH_M7, PATH_M2Q, PATH_M7Q, PATH_QR, PATH_QS = 0, 1, 2, 3, 4, 5, 6
FP64, MIXED = 0, 1

EXPORTED = [
    "cfft_last_error", "cfft_find_optimal_shift", "cfft_count_special_positions",
    "cfft_calc_spectral_asu", "cfft_auto_L", "cfft_select_rep_ops", "cfft_detect_centering",
    "cfft_fill_map", "cfft_plan_create", "cfft_plan_info", "cfft_plan_tables",
    "cfft_update_kernel", "cfft_plan_destroy", "cfft_forward", "cfft_backward",
    "cfft_forward_g0", "cfft_diffuse", "cfft_step", "cfft_sub_to_full", "cfft_full_to_sub",
    "cfft_forward_dev", "cfft_backward_dev", "cfft_diffuse_dev", "cfft_diffuse_noalloc_dev", "cfft_step_dev",
    "cfft_symmetrize_f64_dev", "cfft_dense_create", "cfft_dense_apply", "cfft_dense_step",
    "cfft_dense_destroy",
    "cfft_exp_dev", "cfft_sync", "cfft_step_ptrs_dev", "cfft_profile", "cfft_profile_read",
    "cfft_time_stage_dev", "cfft_sweep_dev", "cfft_plan_set_option",
    "cfft_calc_asu", "cfft_pack_asu", "cfft_block_plan_create", "cfft_block_plan_info",
    "cfft_block_forward", "cfft_block_inverse", "cfft_block_plan_destroy",
    "cfft_prefold_plan_create", "cfft_prefold_plan_info", "cfft_prefold_forward",
    "cfft_prefold_plan_destroy", "cfft_sweep_hist_dev", "cfft_scft_create", "cfft_scft_set_mixing",
    "cfft_scft_set_fields", "cfft_scft_set_params", "cfft_scft_iterate", "cfft_scft_solve", "cfft_scft_get",
    "cfft_scft_propagators", "cfft_scft_destroy", "cfft_scft_create_multi", "cfft_scft_set_params_inst",
    "cfft_scft_set_fields_inst", "cfft_scft_iterate_multi", "cfft_scft_solve_multi", "cfft_scft_get_inst",
    "cfft_scft_copy_state", "cfft_scft_ensemble_create", "cfft_scft_ensemble_iterate",
    "cfft_scft_ensemble_solve",
    "cfft_scft_ensemble_destroy", "cfft_p3c_create", "cfft_p3c_forward", "cfft_p3c_forward_dev",
    "cfft_p3c_destroy", "cfft_p3c_create_strided",
    "cfft_mlp_create", "cfft_mlp_info",
    "cfft_mlp_get_params", "cfft_mlp_set_params", "cfft_mlp_forward", "cfft_mlp_gradient",
    "cfft_mlp_train", "cfft_mlp_destroy",
]

SCFT_W, SCFT_PHI, SCFT_DEV = 0, 1, 2

OPT_PERSISTENT = 1
OPT_FUSE = 2
OPT_GRAPH = 3
OPT_TILE_ROWS = 4
OPT_FP32_TWIDDLES = 5
OPT_FFT_LANES = 6
OPT_PAIR_THREADS = 7
OPT_AXIS_SPLIT = 8

STAGES = ["p1_fold_fftx", "fft_x", "p2_fft_yz", "fft_z", "p3_recon", "p4_irecon_ifft_zy", "ifft_x",
          "ifft_y", "ifft_z", "p5_ifftx_unfold", "qmul", "sweep", "p51_unfold_fold", "p2q_plane_pair"]


class CFFTError(RuntimeError):
    """A non-zero status from the C ABI (the reference's error()/@assert)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class CFFTLibraryMissing(RuntimeError):
    pass


_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)


def _declare(lib):
    sig = {
        "cfft_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "cfft_find_optimal_shift": (C.c_int, [I32, PI32, PI32, I32, PD, PI32]),
        "cfft_count_special_positions": (C.c_int, [I32, PI32, PI32, I32, PI64]),
        "cfft_calc_spectral_asu": (C.c_int, [I32, PI32, PI32, I32, PI32, PI32, I64, PI64]),
        "cfft_auto_L": (C.c_int, [I32, PI32, I32, PI32]),
        "cfft_select_rep_ops": (C.c_int, [I32, PI32, PI32, I32, PI32, PI32]),
        "cfft_detect_centering": (C.c_int, [I32, PI32, PI32, I32, PI32]),
        "cfft_fill_map": (C.c_int, [I32, PI32, PI32, I32, PI32, PI32]),
        "cfft_plan_create": (C.c_int, [PI32, PI32, I32, D, PD, I32, I32, I32, C.POINTER(P)]),
        "cfft_plan_info": (C.c_int, [P, PI32]),
        "cfft_plan_tables": (C.c_int, [P, PI32, PI32, PD, PI32]),
        "cfft_update_kernel": (C.c_int, [P, D, PD]),
        "cfft_plan_destroy": (C.c_int, [P]),
        "cfft_forward": (C.c_int, [P, PD, PD]),
        "cfft_backward": (C.c_int, [P, PD, PD]),
        "cfft_forward_g0": (C.c_int, [P, PD, PD]),
        "cfft_diffuse": (C.c_int, [P, PD]),
        "cfft_step": (C.c_int, [P, PD, PD, D, I32]),
        "cfft_sub_to_full": (C.c_int, [P, PD, PD]),
        "cfft_full_to_sub": (C.c_int, [P, PD, PD]),
        "cfft_forward_dev": (C.c_int, [P, P, P, P]),
        "cfft_backward_dev": (C.c_int, [P, P, P, P]),
        "cfft_diffuse_dev": (C.c_int, [P, P, P]),
        "cfft_diffuse_noalloc_dev": (C.c_int, [P, P, P]),
        "cfft_symmetrize_f64_dev": (C.c_int, [P, P, P, I64, P]),
        "cfft_dense_create": (C.c_int, [I64, PD, P]),
        "cfft_dense_apply": (C.c_int, [P, PD]),
        "cfft_dense_step": (C.c_int, [P, PD, PD, D]),
        "cfft_dense_destroy": (C.c_int, [P]),
        "cfft_step_dev": (C.c_int, [P, P, P, I32, P]),
        "cfft_exp_dev": (C.c_int, [P, P, P, D, I64, P]),
        "cfft_sync": (C.c_int, [P]),
        "cfft_step_ptrs_dev": (C.c_int, [P, P, C.POINTER(P), I32, P]),
        "cfft_profile": (C.c_int, [P, I32]),
        "cfft_profile_read": (C.c_int, [P, PD, PI64]),
        "cfft_time_stage_dev": (C.c_int, [P, I32, P, C.POINTER(P), I32, P, PD]),
        "cfft_sweep_dev": (C.c_int, [P, P, C.POINTER(P), C.POINTER(P), PI32, I32, I32, P]),
        "cfft_plan_set_option": (C.c_int, [P, I32, I64]),
        "cfft_calc_asu": (C.c_int, [I32, PI32, PI32, I32, PI32, PI32, PI32, I64, PI64]),
        "cfft_pack_asu": (C.c_int, [I32, PI32, PI32, I64, PI32, I64, PI64]),
        "cfft_block_plan_create": (C.c_int, [I32, PI32, I64, C.POINTER(P)]),
        "cfft_block_plan_info": (C.c_int, [P, PI64, PI64, PI32]),
        "cfft_block_forward": (C.c_int, [P, PD, PD]),
        "cfft_block_inverse": (C.c_int, [P, PD, PD]),
        "cfft_block_plan_destroy": (C.c_int, [P]),
        "cfft_prefold_plan_create": (C.c_int, [PI32, PI32, I32, C.POINTER(P)]),
        "cfft_prefold_plan_info": (C.c_int, [P, PI32, PI32]),
        "cfft_prefold_forward": (C.c_int, [P, PD, PD]),
        "cfft_prefold_plan_destroy": (C.c_int, [P]),
        "cfft_sweep_hist_dev": (C.c_int, [P, P, C.POINTER(P), C.POINTER(P), PI32, I32, I32, P]),
        "cfft_scft_create": (C.c_int, [P, D, D, I32, I32, C.POINTER(P)]),
        "cfft_scft_set_mixing": (C.c_int, [P, D, D, D]),
        "cfft_scft_set_fields": (C.c_int, [P, PD, PD]),
        "cfft_scft_set_params": (C.c_int, [P, D, D]),
        "cfft_scft_iterate": (C.c_int, [P, PD]),
        "cfft_scft_solve": (C.c_int, [P, D, I32, PI32, PD]),
        "cfft_scft_get": (C.c_int, [P, I32, PD]),
        "cfft_scft_propagators": (C.c_int, [P, C.POINTER(P)]),
        "cfft_scft_destroy": (C.c_int, [P]),
        "cfft_scft_create_multi": (C.c_int, [P, I32, PD, PD, I32, I32, C.POINTER(P)]),
        "cfft_scft_set_params_inst": (C.c_int, [P, I32, D, D]),
        "cfft_scft_set_fields_inst": (C.c_int, [P, I32, PD, PD]),
        "cfft_scft_iterate_multi": (C.c_int, [P, PD]),
        "cfft_scft_solve_multi": (C.c_int, [P, D, I32, PI32, PD]),
        "cfft_scft_get_inst": (C.c_int, [P, I32, I32, PD]),
        "cfft_scft_copy_state": (C.c_int, [P, P]),
        "cfft_scft_ensemble_create": (C.c_int, [C.POINTER(P), I32, C.POINTER(P)]),
        "cfft_scft_ensemble_iterate": (C.c_int, [P, PD]),
        "cfft_scft_ensemble_solve": (C.c_int, [P, D, I32, PI32, PD]),
        "cfft_scft_ensemble_destroy": (C.c_int, [P]),
        "cfft_p3c_create": (C.c_int, [I32, I32, PI32, I32, PI64, C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                      C.POINTER(P)]),
        "cfft_p3c_create_strided": (C.c_int, [I32, I32, I32, PI32, I32, PI64, C.POINTER(P), C.POINTER(P),
                                              C.POINTER(P), C.POINTER(P)]),
        "cfft_p3c_forward": (C.c_int, [P, PD, PD]),
        "cfft_p3c_forward_dev": (C.c_int, [P, P, P, P]),
        "cfft_p3c_destroy": (C.c_int, [P]),
        "cfft_mlp_create": (C.c_int, [I32, PI32, C.c_uint64, C.POINTER(P)]),
        "cfft_mlp_info": (C.c_int, [P, PI64]),
        "cfft_mlp_get_params": (C.c_int, [P, PF]),
        "cfft_mlp_set_params": (C.c_int, [P, PF]),
        "cfft_mlp_forward": (C.c_int, [P, PF, I32, PF]),
        "cfft_mlp_gradient": (C.c_int, [P, PF, PF, I32, PD, PF]),
        "cfft_mlp_train": (C.c_int, [P, PF, PF, I64, I32, I32, C.c_float, C.c_uint64, PD]),
        "cfft_mlp_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load (once) and return the library; raise if it is not built."""
    global _lib
    if _lib is None:
        # PyTorch wheels bundle their own HIP/HSA runtime (same SONAME
        # libamdhip64.so.7).  When both are used in one process, loading torch
        # first makes the library bind to that runtime; loading this library
        # first would load /opt/rocm's runtime beside torch's HSA runtime and
        # torch would then find no GPU.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        if not os.path.exists(LIB_PATH):
            raise CFFTLibraryMissing(
                f"{LIB_PATH} not found: build it with `python crystallographicfft.jl_amd/build.py`")
        l = C.CDLL(LIB_PATH)
        _declare(l)
        _lib = l
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().cfft_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(status: int):
    if status != CFFT_OK:
        raise CFFTError(status, last_error())


def i64p(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PI64)


def i32p(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(PI32)


def f64p(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(PD)
