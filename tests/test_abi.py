"""The C ABI: every function declared in include/radhip.h is exported by the
built library, and the Python layer refuses to run without a GPU (no CPU
fallback).  No compute calls are made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "radhip.h")
LIB = os.path.join(ROOT, "radiacuda_amd", "libradhip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_entry_points():
    names = declared()
    for must in ["rh_model_create", "rh_matrix_assemble", "rh_relax", "rh_field",
                 "rh_current_field", "rh_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        from radiacuda_amd import build
        build.build()
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_no_cpu_fallback_without_gpu():
    L = ctypes.CDLL(LIB)
    L.rh_device_count.restype = ctypes.c_int32
    if L.rh_device_count() > 0:
        pytest.skip("a GPU is visible")
    from radiacuda_amd import _native as N
    from radiacuda_amd import rad
    N._L = None
    with pytest.raises(N.RadHipError):
        N.lib()
    k = rad.ObjRecMag([0, 0, 0], [1, 1, 1], [0, 0, 1])
    with pytest.raises(N.RadHipError):
        rad.Fld(k, "b", [0.5, 0.6, 0.7])
