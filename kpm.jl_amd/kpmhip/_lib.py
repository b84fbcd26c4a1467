"""ctypes binding of libkpm_hip.so (the C ABI declared in include/kpm_hip.h).

The library is the only compute path of this package: if it is missing or no
GPU is visible, every engine call raises instead of falling back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# KPM_HIP_LIB overrides the in-tree build (e.g. a packaged copy of the library)
LIB_PATH = os.environ.get("KPM_HIP_LIB") or os.path.join(_HERE, "libkpm_hip.so")

KPM_OK = 0
KPM_EINVAL = -1
KPM_ENOMEM = -2
KPM_EHIP = -3
KPM_EUNSTABLE = -4
KPM_EUNIMPL = -5

VEC_F64 = 0
VEC_F32 = 1
GEMM_F64 = 0
GEMM_F32 = 1
GEMM_BF16X3 = 2
PARITY = {"NONE": 0, "ODD": 1, "EVEN": 2}
# kpm_ctx_counters indices
COUNT_STEP, COUNT_DUAL, COUNT_DOT, COUNT_COMBINE, COUNT_SYNC = 0, 1, 2, 3, 4
COUNT_N = 8


class ArgumentError(ValueError):
    """Julia's ArgumentError (invalid argument values)."""


class KPMHipError(RuntimeError):
    """A HIP runtime / allocation failure inside libkpm_hip."""


class KPMUnstableError(RuntimeError):
    """Julia's ErrorException from the recurrence stability guard."""


class DomainError(ValueError):
    """Julia's DomainError (an argument outside a function's domain)."""


class QuadratureError(RuntimeError):
    """Julia's ErrorException from the spectral quadrature error contract."""


class KPMHipMissing(RuntimeError):
    """libkpm_hip.so is not built or no GPU is visible."""


_p = C.c_void_p
_i64 = C.c_int64
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)

_SIGS = {
    "kpm_last_error": (C.c_char_p, []),
    "kpm_abi_version": (C.c_int, []),
    "kpm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "kpm_ctx_create": (C.c_int, [C.c_int, C.POINTER(_p)]),
    "kpm_ctx_destroy": (C.c_int, [_p]),
    "kpm_ctx_synchronize": (C.c_int, [_p]),
    "kpm_ctx_set_dual_pass": (C.c_int, [_p, C.c_int]),
    "kpm_ctx_set_bf16_planes": (C.c_int, [_p, C.c_int]),
    "kpm_device_free_memory": (C.c_int, [_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "kpm_device_name": (C.c_int, [_p, C.c_char_p, C.c_size_t]),
    "kpm_csr_from_csc": (C.c_int, [_p, _i64, _pi64, _pi64, _pd, C.c_int, C.POINTER(_p)]),
    "kpm_csr_adjoint_from_csc": (C.c_int, [_p, _i64, _pi64, _pi64, _pd, C.c_int, C.POINTER(_p)]),
    "kpm_csr_from_csr": (C.c_int, [_p, _i64, _pi64, _pi64, _pd, C.c_int, C.POINTER(_p)]),
    "kpm_csr_destroy": (C.c_int, [_p]),
    "kpm_csr_info": (C.c_int, [_p, _pi64, _pi64, C.POINTER(C.c_int)]),
    "kpm_block_create": (C.c_int, [_p, _i64, _i64, C.c_int, C.POINTER(_p)]),
    "kpm_block_destroy": (C.c_int, [_p]),
    "kpm_block_upload": (C.c_int, [_p, _pd]),
    "kpm_block_download": (C.c_int, [_p, _pd]),
    "kpm_block_copy": (C.c_int, [_p, _p]),
    "kpm_block_zero": (C.c_int, [_p]),
    "kpm_block_random_phase": (C.c_int, [_p, C.c_uint64, _i64]),
    "kpm_cheb_step": (C.c_int, [_p, _p, _p, _p, _p]),
    "kpm_spmm": (C.c_int, [_p, _p, C.c_double, _p, C.c_double, _p]),
    "kpm_colwise_dot": (C.c_int, [_p, _pd, _i64, _i64, _p, _p, C.c_double, _i64, C.c_double]),
    "kpm_accum_create": (C.c_int, [_p, _i64, _i64, C.POINTER(_p)]),
    "kpm_accum_destroy": (C.c_int, [_p]),
    "kpm_accum_dot": (C.c_int, [_p, _p, _i64, _p, _p, C.c_double, _i64, C.c_double]),
    "kpm_accum_download": (C.c_int, [_p, _p, _pd]),
    "kpm_ctx_set_merged_limit": (C.c_int, [_p, _i64]),
    "kpm_ctx_counters": (C.c_int, [_p, _pi64, C.c_int, C.c_int]),
    "kpm_batch_graph_builds": (C.c_int, [_p, _pi64]),
    "kpm_1d": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd, C.c_int]),
    "kpm_1d_lr": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd, _pd, C.c_int]),
    "kpm_1d_current": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _pd, _pd, C.c_int]),
    "kpm_1d_block": (C.c_int, [_p, _p, _i64, _p, _pd]),
    "kpm_1d_avg": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd, C.c_int]),
    "kpm_1d_block_avg": (C.c_int, [_p, _p, _i64, _p, _pd]),
    "kpm_1d_block_noalloc": (C.c_int, [_p, _p, _i64, _p, _p, _p, _p]),
    "kpm_1d_lr_block": (C.c_int, [_p, _p, _i64, _p, _p, _pd]),
    "kpm_2d_blocking": (C.c_int, [_p, _i64, _i64, _i64, _i64, _i64, _i64, _pi64, _pi64, _pi64]),
    "kpm_2d": (C.c_int, [_p, _p, _p, _p, _i64, _i64, _i64, _pd, _pd, _i64, _i64, C.c_int,
                         C.c_int, C.c_int, C.c_int, _pd]),
    "kpm_3d": (C.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _i64, _pd, _pd, _i64, _i64, C.c_int,
                         C.c_int, _pd]),
    "kpm_cheb_action": (C.c_int, [_p, _p, _i64, _i64, _i64, _i64, _pd, _pd, _pd, C.c_int, _i64]),
    "kpm_cheb_action_block": (C.c_int, [_p, _p, _i64, _i64, _p, _pd, _p, _i64]),
    "kpm_spectral_edges": (C.c_int, [_p, _p, _i64, C.c_double, C.c_uint64, _pd, _pd, _pi64]),
    "kpm_cheb_lin_trans": (C.c_int, [_p, _i64, _pd, _i64, _pi64, _pd, _pd]),
    "kpm_kubo_bastin_nodes": (C.c_int, [_p, _i64, _pd, _i64, _pd, _pd, _pd]),
    "kpm_table3_create": (C.c_int, [_p, _i64, _pd, C.POINTER(C.c_void_p)]),
    "kpm_table3_destroy": (C.c_int, [_p]),
    "kpm_trilinear_forms": (C.c_int, [_p, _i64, _pd, _pd, _pd, _pd]),
    "kpm_bench_trilinear": (C.c_int, [_p, _i64, _i64, C.POINTER(C.c_float)]),
    "kpm_trilinear_forms_rec": (C.c_int, [_p, _i64, _pd, _pd, _pd, _pd]),
    "kpm_table3_create_weighted": (C.c_int, [_p, _i64, _pd, _pd, C.POINTER(C.c_void_p)]),
    "kpm_bond_gradient": (C.c_int, [_i64, _pi64, _pi64, _pd, _pd, C.c_double, _pi64, _pi64, _pd,
                                    _pi64]),
    "kpm_1d_adjoint": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd, _pd, C.c_int, _i64]),
    "kpm_1d_adjoint_forward": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd, C.c_int, _i64]),
    "kpm_1d_adjoint_reverse": (C.c_int, [_p, _p, _pd, _pd]),
    "kpm_2d_adjoint": (C.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _i64, _pd, _pd, _pd, _pd, _pd,
                                 _pd]),
    "kpm_batch_create": (C.c_int, [_p, C.c_int, _pi64, _pi64, _pi64, _pd, C.c_int, C.POINTER(_p)]),
    "kpm_batch_destroy": (C.c_int, [_p]),
    "kpm_batch_1d": (C.c_int, [_p, _p, _i64, _i64, _pd, _pd, C.c_int, C.c_int]),
    "kpm_batch_info": (C.c_int, [_p, _pi64, _pi64, C.POINTER(C.c_int)]),
    "kpm_batch_upload": (C.c_int, [_p, _p, _i64, _pd, C.c_int]),
    "kpm_batch_run": (C.c_int, [_p, _p, _i64, _pd, C.c_int]),
    "kpm_batch_autotune": (C.c_int, [_p, _p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float)]),
    "kpm_bench_step": (C.c_int, [_p, _p, _p, _p, _i64, C.POINTER(C.c_float)]),
    "kpm_bench_step_mode": (C.c_int, [_p, _p, _p, _p, _i64, C.c_int, C.POINTER(C.c_float)]),
    "kpm_bench_copy": (C.c_int, [_p, _p, _p, _i64, C.POINTER(C.c_float)]),
    "kpm_bench_contract": (C.c_int, [_p, _i64, _i64, _i64, _i64, C.c_int, C.c_int, _i64,
                                     C.POINTER(C.c_float)]),
    "kpm_event_record": (C.c_int, [_p, C.c_int]),
    "kpm_event_elapsed": (C.c_int, [_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "kpm_autotune_step": (C.c_int, [_p, _p, _p, _p, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "kpm_tune_set": (C.c_int, [_p, _p, _i64, C.c_int, C.c_int, C.c_int]),
    "kpm_tune_set_chunked": (C.c_int, [_p, _p, _i64, C.c_int, C.c_int, C.c_int, C.c_int]),
    "kpm_tune_get": (C.c_int, [_p, _p, _i64, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "kpm_tune_get_mode": (C.c_int, [_p, _p, _i64, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
}

EXPORTED = tuple(_SIGS)

_lib = None
_lock = threading.Lock()


def load_library(path: str | None = None) -> C.CDLL:
    """Load libkpm_hip.so (no GPU needed to load it)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise KPMHipMissing(
                f"{p} is not built; build it with `python kpm.jl_amd/build_hip.py` "
                "(there is no CPU fallback)")
        lib = C.CDLL(p)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    msg = load_library().kpm_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(rc: int, what: str = "") -> None:
    """Map a libkpm_hip status to the reference's exception types."""
    if rc == KPM_OK:
        return
    msg = last_error() or what
    if rc == KPM_EINVAL:
        raise ArgumentError(msg)
    if rc == KPM_EUNIMPL:
        raise NotImplementedError(msg)
    if rc == KPM_ENOMEM:
        raise MemoryError(msg)
    if rc == KPM_EUNSTABLE:
        raise KPMUnstableError(msg)
    raise KPMHipError(f"{what}: {msg} (status {rc})")


def dptr(a: np.ndarray):
    """double* of a C- or F-contiguous float64/complex128 array (complex data
    is interleaved re, im)."""
    if not (a.flags.f_contiguous or a.flags.c_contiguous):
        raise ValueError("array must be contiguous")
    return a.ctypes.data_as(_pd)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_pi64)
