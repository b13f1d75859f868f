"""ctypes binding of the C ABI in include/pointint_hip.h (libpointint_hip.so).

The library is built in-tree by ``pointint_amd.build_lib`` (hipcc, gfx950).
There is no fallback: if the library is missing or fails to load, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpointint_hip.so")

PII_OK = 0
PII_ERR_INVALID = 1
PII_ERR_HIP = 2
PII_ERR_NOMEM = 3
PII_ERR_UNSUPPORTED = 4
PII_ERR_BUDGET = 5
PII_ERR_IO = 6
PII_ERR_BUSY = 7

GEOM_TRIANGLE, GEOM_DISK, GEOM_GAUSSIAN = 0, 1, 2
DOF_AFFINE, DOF_TRIMESH = 0, 1
FF_FP64, FF_FAST = 0, 1
GRAM_FP64, GRAM_BF16X3, GRAM_FP16X3 = 0, 1, 2
FLAG_VOLUME, FLAG_GRADIENT, FLAG_HESSIAN = 1, 2, 4

# Every exported symbol and its (argtypes, restype).  The CPU tests check that
# the built library exports exactly what include/pointint_hip.h declares.
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_S = C.c_char_p
SIGNATURES = {
    "pii_last_error": ([], _S),
    "pii_version": ([], _S),
    "pii_debug_fail_allocation": ([_I], _I),
    "pii_gauss_legendre": ([_I, _D, _D, _P, _P], _I),
    "pii_lebedev_load": ([_S, _P, _P, _I, _P], _I),
    "pii_kgrid_build": ([_P, _P, _I, _I, _P, _P, _P], _I),
    "pii_context_create": ([_I, _P], _I),
    "pii_context_destroy": ([_P], _I),
    "pii_context_synchronize": ([_P], _I),
    "pii_context_set_profiling": ([_P, _I], _I),
    "pii_context_timings": ([_P, _P, _I], _I),
    "pii_context_set_series": ([_P, _I], _I),
    "pii_context_work": ([_P, _P, _I], _I),
    "pii_context_set_caller_stream": ([_P, _P], _I),
    "pii_volume_matrix_hessian": ([_P, _I, _P, _P, _P, _P, _P, _I, _P, _P], _I),
    "pii_volume_matrix_affine_derivatives": ([_P, _I, _P, _P, _P, _I, _P, _P, _P], _I),
    "pii_volume_matrix_derivatives": ([_P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P], _I),
    "pii_measure_issue_rate": ([_P, _I, _P], _I),
    "pii_kgrid_create": ([_P, _I, _P, _P, _P, _P], _I),
    "pii_kgrid_destroy": ([_P], _I),
    "pii_kgrid_info": ([_P, _P, _P, _P], _I),
    "pii_geometry_create_mesh": ([_P, _P, _I, _P, _I, _P], _I),
    "pii_geometry_create_tris": ([_P, _P, _P, _I, _P], _I),
    "pii_geometry_create_disks": ([_P, _P, _I, _P], _I),
    "pii_geometry_create_gaussians": ([_P, _P, _I, _P], _I),
    "pii_geometry_destroy": ([_P], _I),
    "pii_geometry_info": ([_P, _P, _P, _P], _I),
    "pii_geometry_cache_moments": ([_P, _I, _P], _I),
    "pii_volume_matrix_device": ([_P, _I, _P, _P, _P, _P, _P, _I, _I, _P], _I),
    "pii_volume_matrix_vjp_device": ([_P, _I, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "pii_form_factor": ([_P, _P, _I, _P, _I, _P, _I, _P], _I),
    "pii_form_factor_device": ([_P, _P, _I, _P, _I, _P, _I, _P], _I),
    "pii_form_factor_device_noalloc": ([_P, _P, _I, _P, _I, _P, _I, _P], _I),
    "pii_form_factor_jacobian": ([_P, _P, _I, _P, _I, _P, _P], _I),
    "pii_form_factor_jacobian_device": ([_P, _P, _I, _P, _I, _P, _P], _I),
    "pii_volume_matrix": ([_P, _I, _P, _P, _P, _P, _P, _I, _I, _P], _I),
    "pii_gram": ([_P, _P, _I, _P, _I, _P], _I),
    "pii_volume_matrix_vjp": ([_P, _I, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "pii_clock_seconds": ([_P], _I),
    "pii_server_create": ([_P, _P, _I, _D, _I, _P], _I),
    "pii_server_destroy": ([_P], _I),
    "pii_server_set_latency_target": ([_P, _D], _I),
    "pii_server_set_graphs": ([_P, _I], _I),
    "pii_server_submit": ([_P, _I, _P, _P, _P], _I),
    "pii_server_submit_at": ([_P, _I, _P, _P, _D, _P], _I),
    "pii_server_wait": ([_P, C.c_int64, _D, _P, _P], _I),
    "pii_server_wait_times": ([_P, C.c_int64, _D, _P, _P], _I),
    "pii_server_stats": ([_P, _P, _I], _I),
}

_lib = None


class PiiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pii status {code}] {msg}")
        self.code = code


def _share_torch_hip_runtime():
    """PyTorch-ROCm wheels bundle their own libamdhip64.so.7.  When torch is
    installed it is imported before the library is opened, so the library's
    NEEDED libamdhip64.so.7 resolves to the copy torch already loaded and the
    process holds one HIP runtime (two runtimes in one process leave the second
    one without devices)."""
    if "torch" in sys.modules:
        return
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load(path: str = LIB_PATH):
    """Load libpointint_hip.so (raises if it is missing or incomplete)."""
    global _lib
    if _lib is not None:
        return _lib
    _share_torch_hip_runtime()
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m pointint_amd.build_lib` "
            "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != PII_OK:
        msg = load().pii_last_error()
        raise PiiError(code, msg.decode() if msg else "")


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)
