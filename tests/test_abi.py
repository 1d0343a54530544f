"""The C ABI library loads and exports every function include/wolfd2_amd.h declares
(no compute call: runs without a GPU)."""
import ctypes as C
import os

from wolfd2_amd import _lib


def test_library_exports_every_declared_function():
    names = _lib.header_functions()
    assert len(names) >= 20
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_has_gfx950_code_object():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_oracle_is_not_linked_into_the_product():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"wd2o_" not in data
    assert os.path.basename(_lib.LIB_PATH) == "libwolfd2_amd.so"
