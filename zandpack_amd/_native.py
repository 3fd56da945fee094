"""ctypes binding of the C ABI declared in include/zandhip.h.

The library is the product path: there is no CPU fallback.  Loading fails with
an ImportError naming the missing file when ``libzandhip.so`` was not built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzandhip.so")

ZH_FP64, ZH_MIXED = 0, 1
ZH_TIME_REFERENCE, ZH_TIME_CORRECT = 0, 1
ZH_ELIMIT = -6
WAVE_KINDS = {"const": 0, "cos_window": 1, "sin_fermi": 2, "gauss_sin": 3}
PULSE_KINDS = {"none": 0, "step": 1, "gauss": 2, "ramp": 3, "sin": 4, "box": 5, "generic": 6,
               "stairs": 7, "spline": 8}
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ESTATE", -5: "ENONFINITE", -6: "ELIMIT"}


class ZHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zandhip {ERRORS.get(code, code)}: {msg}")
        self.code = code


class zh_dims(C.Structure):
    _fields_ = [("NK", C.c_int), ("n", C.c_int), ("NL", C.c_int), ("NX", C.c_int),
                ("NOT", C.c_int), ("precision", C.c_int), ("hbar", C.c_double),
                ("device", C.c_int)]


class zh_precond(C.Structure):
    _fields_ = [("evecs", C.POINTER(C.c_double)), ("evals", C.POINTER(C.c_double)),
                ("ncr", C.c_int), ("corb", C.POINTER(C.c_int)), ("ngroups", C.c_int),
                ("group", C.POINTER(C.c_int)), ("K", C.POINTER(C.c_double))]


class zh_attempt_info(C.Structure):
    _fields_ = [("t0", C.c_double), ("h", C.c_double), ("eps", C.c_double),
                ("force_accept", C.c_int), ("te", C.c_double), ("h_next", C.c_double),
                ("accepted", C.c_int)]


_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)

# name -> (restype, argtypes); must match include/zandhip.h
SIGNATURES = {
    "zh_abi_version": (C.c_int, []),
    "zh_last_error": (C.c_char_p, []),
    "zh_problem_create": (C.c_int, [C.POINTER(zh_dims), C.POINTER(C.c_void_p)]),
    "zh_problem_create_instances": (C.c_int, [C.POINTER(zh_dims), C.c_int,
                                               C.POINTER(C.c_void_p)]),
    "zh_instances": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_instance_pulse": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp]),
    "zh_attempt_instances": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info), _dp]),
    "zh_launch_instances": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info)]),
    "zh_collect_instances": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info), _dp]),
    "zh_problem_destroy": (None, [C.c_void_p]),
    "zh_set_constants": (C.c_int, [C.c_void_p] + [_dp] * 8),
    "zh_set_hubbard": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "zh_set_drive": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _dp, _dp,
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int), _dp]),
    "zh_set_pulse": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp]),
    "zh_state_put": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "zh_state_get": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "zh_sigma_get": (C.c_int, [C.c_void_p, _dp]),
    "zh_rhs": (C.c_int, [C.c_void_p] + [_dp] * 9),
    "zh_attempt": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info), _dp]),
    "zh_attempt_launch": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info)]),
    "zh_attempt_collect": (C.c_int, [C.c_void_p, C.POINTER(zh_attempt_info), _dp]),
    "zh_attempt_ready": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_batch_attempt": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(zh_attempt_info),
                                   C.POINTER(_dp), C.POINTER(C.c_int)]),
    "zh_problem_bytes": (C.c_int, [C.c_void_p, _lp]),
    "zh_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "zh_set_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_propagate": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_int, C.c_long, C.c_long, _dp, _dp, _dp, _lp, _lp, _dp]),
    "zh_set_layout": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_layout_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_set_sparse": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_steady_state": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int,
                                  C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "zh_set_dissipator": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                    _dp]),
    "zh_set_tolerance": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "zh_set_overlap": (C.c_int, [C.c_void_p, _dp]),
    "zh_pi_get": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "zh_set_pulse_ex": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _dp, _dp]),
    "zh_set_pulse_spline": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "zh_contour_density": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(C.c_int), _dp, _dp, C.c_int]),
    "zh_contour_density_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                             C.POINTER(C.c_int), _dp, _dp, C.c_int]),
    "zh_window_density": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                    C.POINTER(C.c_int), _dp, _dp, _dp, C.c_int]),
    "zh_hamiltonian_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_set_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_set_omgxi": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_set_psi_kernel": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_psi_kernel_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_set_fsal": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_fsal_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_int)]),
    "zh_omgxi_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_set_antihermitian": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_antihermitian_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "zh_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_set_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "zh_timing": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, _dp, _lp]),
    "zh_selftest_gemm": (C.c_int, [C.c_int] * 6 + [_dp] * 3),
}

_lib = None


def lib():
    """Load libzandhip.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: build it with "
                              "`python -m zandpack_amd.build` (no CPU fallback exists)")
        L = C.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.zh_abi_version() != 1:
            raise ImportError("libzandhip.so ABI version mismatch")
        _lib = L
    return _lib


def check(code):
    if code != 0:
        raise ZHError(code, lib().zh_last_error().decode())
    return code


def ptr(a):
    """double* view of a C-contiguous float64/complex128 array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128), a.dtype
    return a.ctypes.data_as(_dp)


def c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)
