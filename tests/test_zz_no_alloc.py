"""wd2_alt_trid_lu_noalloc: AltTridLU (momentum.f:1307-1339) on device arrays
with its scratch in a static device array (no allocation by the library).

Bar: bit for bit equal to the allocating entry point wd2_alt_trid_lu and to
the oracle's AltTridLU, on calls in a row that shrink (the static scratch
keeps the previous call's contents), including a chain that takes the
sequential fallback; sizes beyond the scratch are refused.
"""
import ctypes as C

import numpy as np
import pytest

from wolfd2_amd import api
from wolfd2_amd._lib import WD2_ERR_ARG, lib

MAX_ROWS = 266240   # WD2_TRID_NOALLOC_MAX_ROWS of include/wolfd2_amd.h


def _chain(n, rng):
    a = np.zeros((n, 3))
    a[:, 0] = rng.uniform(-1, 0, n)
    a[:, 2] = rng.uniform(-1, 0, n)
    a[:, 1] = (np.abs(a[:, 0]) + np.abs(a[:, 2])) * 1.3 + 0.05
    return a, rng.standard_normal(n)


def _laplacian(n, rng):
    a = np.zeros((n, 3))
    a[:, 0], a[:, 1], a[:, 2] = -1.0, 2.0, -1.0
    return a, rng.standard_normal(n)


def _hip():
    """The HIP runtime the library itself links (so the test allocates its
    device arrays in the same runtime, whatever else the process loaded)."""
    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _noalloc(a, b):
    """Caller-owned device arrays (the test allocates them; the entry point
    does not), the call on a stream of the caller's, then the solution back."""
    h = _hip()
    H2D, D2H = 1, 2
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    x = np.array(b, dtype=np.float64)
    da, db, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(da), a.nbytes) == 0 and h.hipMalloc(C.byref(db), x.nbytes) == 0
    assert h.hipStreamCreate(C.byref(st)) == 0
    try:
        assert h.hipMemcpy(da, a.ctypes.data, a.nbytes, H2D) == 0
        assert h.hipMemcpy(db, x.ctypes.data, x.nbytes, H2D) == 0
        rc = lib().wd2_alt_trid_lu_noalloc(x.size, da, db, st)
        assert rc == 0, lib().wd2_last_error().decode()
        assert h.hipStreamSynchronize(st) == 0
        assert h.hipMemcpy(x.ctypes.data, db, x.nbytes, D2H) == 0
    finally:
        h.hipStreamDestroy(st)
        h.hipFree(da)
        h.hipFree(db)
    return x


@pytest.mark.gpu
def test_noalloc_trid_equals_allocating_entry_point(oracle):
    rng = np.random.default_rng(20261203)
    cases = [("random", MAX_ROWS, _chain), ("random", 100_003, _chain), ("laplacian", 5000, _laplacian),
             ("random", 3, _chain)]
    for kind, n, make in cases:
        print(f"noalloc AltTridLU: {kind} chain n={n}", flush=True)
        a, b = make(n, rng)
        x = _noalloc(a, b)
        ref = api.alt_trid_lu(a, b)
        _, x_oracle = oracle.alt_trid_lu(a, b)
        assert np.array_equal(ref, x_oracle)
        assert np.array_equal(x, ref), f"{kind} n={n}"


def test_noalloc_trid_refuses_sizes_beyond_its_scratch():
    # the size check comes before any device work, so this runs without a GPU
    rc = lib().wd2_alt_trid_lu_noalloc(MAX_ROWS + 1, C.c_void_p(1), C.c_void_p(1), None)
    assert rc == WD2_ERR_ARG
    assert b"WD2_TRID_NOALLOC_MAX_ROWS" in lib().wd2_last_error()
    assert lib().wd2_alt_trid_lu_noalloc(1, C.c_void_p(1), C.c_void_p(1), None) == WD2_ERR_ARG
