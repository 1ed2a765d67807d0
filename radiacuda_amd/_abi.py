"""ctypes mirror of the C ABI structs in include/radhip.h."""
import ctypes

import numpy as np

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)

RH_OK, RH_ERR_ARG, RH_ERR_HIP, RH_ERR_NOMEM, RH_ERR_UNSUPPORTED, RH_ERR_GEOMETRY, \
    RH_ERR_NOTCONV = range(7)
RH_FP64, RH_MIXED = 0, 1
RH_FALLBACK_CPU, RH_FALLBACK_GPU_STREAMING, RH_FALLBACK_BREAK = 0, 1, 2
RH_STREAM_BY_POLICY = -1
MAT_STRIDE = 16


class ModelDesc(ctypes.Structure):
    _fields_ = [("n_elem", ctypes.c_int64), ("elem_type", _i32), ("rec_center", _d),
                ("rec_dims", _d), ("vert_off", _i64), ("verts", _d), ("face_off", _i64),
                ("face_len", _i32), ("face_idx", _i32), ("copy_off", _i64), ("copy_m", _d),
                ("copy_v", _d), ("copy_s", _d)]


class RelaxOpts(ctypes.Structure):
    _fields_ = [("prec", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("keep_m", ctypes.c_int32), ("omega0", ctypes.c_double),
                ("method", ctypes.c_int32), ("refine", ctypes.c_int32),
                ("krylov_m", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MatrixOpts(ctypes.Structure):
    _fields_ = [("device_bytes", ctypes.c_int64), ("resident_fraction", ctypes.c_double),
                ("allow_streaming", ctypes.c_int32), ("fp64_integrals", ctypes.c_int32),
                ("dense_rows", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RelaxResult(ctypes.Structure):
    _fields_ = [("misfit", ctypes.c_double), ("max_m", ctypes.c_double),
                ("max_h", ctypes.c_double), ("iters", ctypes.c_int32),
                ("matvecs", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("refined", ctypes.c_int32), ("newton", ctypes.c_int32),
                ("krylov", ctypes.c_int32)]


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def model_desc(packed):
    """Build a ModelDesc from ``objects.pack_elements`` output.

    Returns (desc, keepalive); keep ``keepalive`` referenced while desc is used."""
    n = len(packed["type"])
    arrs = dict(
        elem_type=_c(packed["type"], np.int32), rec_center=_c(packed["rec_c"], np.float64),
        rec_dims=_c(packed["rec_d"], np.float64), vert_off=_c(packed["vert_off"], np.int64),
        verts=_c(packed["verts"], np.float64), face_off=_c(packed["face_off"], np.int64),
        face_len=_c(packed["face_len"], np.int32), face_idx=_c(packed["face_idx"], np.int32),
        copy_off=_c(packed["copy_off"], np.int64), copy_m=_c(packed["cM"], np.float64),
        copy_v=_c(packed["cV"], np.float64), copy_s=_c(packed["cS"], np.float64))
    # non-empty buffers so that pointers are valid
    for k, a in list(arrs.items()):
        if a.size == 0:
            arrs[k] = np.zeros(1, dtype=a.dtype)
    ptr = {np.dtype(np.float64): _d, np.dtype(np.int32): _i32, np.dtype(np.int64): _i64}
    desc = ModelDesc(n, *[arrs[f].ctypes.data_as(ptr[arrs[f].dtype]) for f, _ in
                         ModelDesc._fields_[1:]])
    return desc, arrs
