"""Loader of libwolfd2_amd.so (the HIP path) through its C ABI (include/wolfd2_amd.h).

There is no CPU fallback: if the library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwolfd2_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "wolfd2_amd.h")

WD2_OK, WD2_ERR_FLAG, WD2_ERR_ARG, WD2_ERR_HIP, WD2_ERR_UNSUPPORTED = 0, -1, -2, -3, -4


class Wd2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wolfd2_amd error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """wd2_params of include/wolfd2_amd.h."""
    _fields_ = [("lCartesGrid", C.c_int), ("mqiter", C.c_int), ("nPpeSolver", C.c_int), ("msorit", C.c_int),
                ("qtol", C.c_double), ("sortol", C.c_double), ("sorrel", C.c_double),
                ("dk", C.c_double), ("re", C.c_double), ("fr", C.c_double)]


class StepOptions(C.Structure):
    """wd2_step_options of include/wolfd2_amd.h."""
    _fields_ = [("nthermen", C.c_int), ("neqstate", C.c_int), ("nmeiter", C.c_int), ("nfiltu", C.c_int),
                ("nfiltv", C.c_int), ("nfiltt", C.c_int), ("dmeittol", C.c_double), ("fpu", C.c_double),
                ("fpv", C.c_double), ("fpt", C.c_double), ("pe", C.c_double), ("uref", C.c_double),
                ("densref", C.c_double), ("tmax", C.c_double), ("tref", C.c_double), ("rconst", C.c_double)]


class AdjStats(C.Structure):
    """wd2_adj_stats of include/wolfd2_amd.h."""
    _fields_ = [("steps", C.c_int), ("adj_iters_max", C.c_int), ("adj_iters_mean", C.c_double),
                ("ppe_cycles_max", C.c_int), ("ppe_cycles_mean", C.c_double), ("fwd_unconverged", C.c_int),
                ("adj_unconverged", C.c_int), ("ms_forward", C.c_double), ("ms_recompute", C.c_double),
                ("ms_reverse", C.c_double)]


_lib = None


def header_functions() -> list[str]:
    """Names of the functions include/wolfd2_amd.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wd2_[a-z0-9_]+)\s*\(", txt)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.wd2_last_error.restype = C.c_char_p
        L.wd2_ctx_stream.restype = C.c_void_p
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != WD2_OK:
        raise Wd2Error(code, lib().wd2_last_error().decode())


def dp(a: np.ndarray):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "fields must be C-contiguous float64 [j, i] arrays"
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ip(a: np.ndarray):
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_int))


def fp(a: np.ndarray):
    """Fortran-ordered float64 table pointer."""
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_double))
