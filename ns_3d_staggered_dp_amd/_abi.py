"""ctypes binding of libns_hip.so (the C ABI declared in include/ns_hip.h).

The library is loaded from the package directory only; if it is missing the
import of the product path fails loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libns_hip.so")

NS_OK = 0
NS_ERR_INVALID = -1
NS_ERR_UNSUPPORTED = -2
NS_ERR_HIP = -3
NS_ERR_NONFINITE = -4
NS_ERR_NOMEM = -5

_ERRS = {
    NS_ERR_INVALID: "invalid argument",
    NS_ERR_UNSUPPORTED: "unsupported configuration",
    NS_ERR_HIP: "HIP runtime error",
    NS_ERR_NONFINITE: "non-finite values",
    NS_ERR_NOMEM: "out of device memory",
}


class NSError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed: {code} ({_ERRS.get(code, 'unknown')})")
        self.code = code


P = ctypes.c_void_p
D = ctypes.c_double
I = ctypes.c_int


class SubBox(ctypes.Structure):
    """ns_subbox: 1-based origin and interior size of a re-simulation box"""
    _fields_ = [("x0", ctypes.c_int), ("y0", ctypes.c_int), ("z0", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "u_star": 4, "v_star": 5, "w_star": 6, "dp": 7, "hx": 8, "hy": 9, "hz": 10}

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "ns_create": [P, ctypes.POINTER(P)],
    "ns_destroy": [P],
    "ns_plane_set_length": [I, I, I],
    "ns_upload_state": [P, P, P, P, P, P, P, P, I],
    "ns_download_state": [P, P, P, P, P, P, P, P],
    "ns_download_work": [P, P, P, P, P],
    "ns_synchronize": [P],
    "ns_state_pointers": [P, P],
    "ns_state_pitch": [P, P],
    "ns_set_bc_planes": [P, P],
    "ns_get_bc_planes": [P, P],
    "ns_bc_interp": [P, P, P, D],
    "ns_step": [P, I],
    "ns_step_profiled": [P, I, P],
    "ns_resim_advance": [P, P, I, D, D, I],
    "ns_capture_begin": [P, P, P, I, I],
    "ns_capture_end": [P, ctypes.POINTER(I)],
    "ns_max_div": [P, ctypes.POINTER(D)],
    "ns_check_finite": [P],
    "ns_stream_copy": [P, P, P, ctypes.c_long],
    "ns_error_metrics": [P, P, P, P, P, ctypes.POINTER(D), ctypes.POINTER(D)],
    "ns_error_metrics_noalloc": [P, P, P, P, P, ctypes.POINTER(D), ctypes.POINTER(D)],
    "ns_resim_advance_spline": [P, P, I, D, D, I],
    "ns_bc_interp_spline": [P, P, I, D, D],
    "ns_spline_weights": [P, I, D, P],
    "ns_set_dt": [P, D],
    "ns_extract_subbox": [P, P, I, P],
    "ns_adjoint_set_state": [P, P, P, P, P, P, P, P],
    "ns_adjoint_get_state": [P, P, P, P, P, P, P, P],
    "ns_adjoint_step": [P, P, P, P, I, P],
    "ns_batch_create": [P, I, P],
    "ns_batch_destroy": [P],
    "ns_batch_step": [P, I],
    "ns_batch_tune": [P, I, P],
    "ns_vel_bc_staggered_adj": [P, P, P, P, P, I, I, I],
    "ns_vel_bc_staggered_cn2_adj": [P, P, P, P, P, I, I],
    "ns_adi_sweep_adj": [P, P, P, P, I],
    "ns_divergence_adj": [P, P, D, P, P, P],
    "ns_pr_bc_staggered_modify_rhs_adj": [P, P, P],
    "ns_poisson_fft_adj": [P, P, P],
    "ns_dct_poisson_solver_adj": [P, P, P],
    "ns_pr_bc_staggered_adj": [P, P, P],
    "ns_project_adj": [P, P, P, P, P, D, I, P, P, P, P],
    "ns_history_save": [P],
    "ns_history_restore": [P],
    "ns_cal_conv": [P, P, P, P, I, I, I, D, D, D, P, P, P],
    "ns_cal_diff": [P, P, P, P, I, I, I, D, D, D, P, P, P],
    "ns_vel_bc_staggered": [P, P, P, P, P, I, I, I],
    "ns_vel_bc_staggered_cn2": [P, P, P, P, P, I, I],
    "ns_adi_sweep": [P, P, P, P, I],
    "ns_get_velpr_bc": [P, P, P, P, P, D, I, I, P, P],
    "ns_get_pr_bc": [P, P, P],
    "ns_divergence": [P, P, P, P, D, P],
    "ns_pr_bc_staggered_modify_rhs": [P, P, P],
    "ns_poisson_fft": [P, P, P],
    "ns_dct_poisson_solver": [P, P, D],
    "ns_pr_bc_staggered": [P, P, P],
    "ns_project": [P, P, P, P, P, D, I, P, P, P, P],
}
_RESTYPE = {"ns_plane_set_length": ctypes.c_long}

_lib = None


def lib():
    """Load libns_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # PyTorch-ROCm bundles its own libamdhip64.so with the same soname
        # (libamdhip64.so.7) as /opt/rocm's.  Loading torch first makes this
        # library bind to that copy, so a process holds one HIP runtime; loading
        # ours first and torch later would start a second runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m ns_3d_staggered_dp_amd.build`"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if rc != NS_OK:
        raise NSError(name, rc)
    return rc


def declared_symbols(header: str) -> list[str]:
    """Entry points declared in a header of include/ (for export checks)."""
    import re

    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|void)\s+\**\s*(nsh?5?_\w+)\s*\(", txt, flags=re.M)))
