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


def test_fallback_policy_through_the_c_abi():
    """rh_set_fallback / rh_get_fallback / rh_last_backend through ctypes alone
    (no device needed): the reference's policy values and its rejection of
    anything else leaving the policy alone (RadUtiGpuFallback,
    radentry.cpp:934-944; tests/test_gpu_fallback_policy.py:51-63)."""
    L = ctypes.CDLL(LIB)
    L.rh_set_fallback.argtypes = [ctypes.c_int32]
    L.rh_set_fallback.restype = ctypes.c_int
    L.rh_get_fallback.restype = ctypes.c_int32
    L.rh_last_backend.argtypes = [ctypes.c_int32]
    L.rh_last_backend.restype = ctypes.c_int32
    L.rh_last_error.restype = ctypes.c_char_p
    start = L.rh_get_fallback()
    try:
        for mode in (2, 1, 0):
            assert L.rh_set_fallback(mode) == 0
            assert L.rh_get_fallback() == mode
        for bad in (-1, 3, 100):
            assert L.rh_set_fallback(bad) == 1  # RH_ERR_ARG
            assert b"fallback mode" in L.rh_last_error()
            assert L.rh_get_fallback() == 0
        # nothing assembled or evaluated in this process yet by this library copy
        assert L.rh_last_backend(7) == -1
        assert L.rh_last_backend(0) in (-1, 1, 2) and L.rh_last_backend(1) in (-1, 1)
    finally:
        L.rh_set_fallback(start)


def test_rad_policy_names_follow_the_reference():
    """UtiGpuFallback maps the reference's exact names onto the C ABI and
    rejects others with its 'expected ...' message (radpy.cpp:2919-2945)."""
    from radiacuda_amd import _native as N
    from radiacuda_amd import rad
    assert rad.UtiGpuFallback() == "cpu"
    for mode in ("break", "gpu_streaming", "cpu"):
        assert rad.UtiGpuFallback(mode) == mode and rad.UtiGpuFallback() == mode
        assert N.get_fallback() == N.FALLBACK_NAMES.index(mode)
    for bad in ("CPU", "stream", "gpu-streaming", "", 1):
        with pytest.raises(RuntimeError, match="expected"):
            rad.UtiGpuFallback(bad)
    assert rad.UtiGpuFallback() == "cpu"


def test_fld_precision_arguments():
    """Fld's precision argument as the reference parses it
    (radpy.cpp:2402-2415): unknown strings raise before any device work."""
    from radiacuda_amd import rad
    k = rad.ObjRecMag([0, 0, 0], [1, 1, 1], [0, 0, 1])
    for bad in ("half", "Single", "DOUBLE", "fp16"):
        with pytest.raises(RuntimeError, match="Incorrect function arguments"):
            rad.Fld(k, "b", [0.5, 0.6, 0.7], precision=bad)
