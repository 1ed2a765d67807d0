"""ctypes binding of libradhip.so (the C ABI in include/radhip.h).

The library is built in-tree by ``radiacuda_amd.build``.  There is no CPU
fallback: if the library or a GPU is missing every entry point raises.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import RH_FP64, RH_MIXED, RH_OK  # noqa: F401  (re-exported)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libradhip.so")
_L = None

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class RadHipError(RuntimeError):
    pass


_RAW = None


def lib():
    """The library, after checking that a HIP device is visible."""
    global _L
    if _L is not None:
        return _L
    L = _load()
    if L.rh_device_count() < 1:
        raise RadHipError("no HIP device visible; the HIP path has no CPU fallback")
    _L = L
    return L


def _load():
    """Loads and binds the library without touching a device (the process-wide
    policy and diagnostics calls need none)."""
    global _RAW
    if _RAW is not None:
        return _RAW
    if not os.path.exists(LIB_PATH):
        raise RadHipError("libradhip.so is not built (run: python -m radiacuda_amd.build); "
                          "the HIP path has no CPU fallback")
    try:
        # One HIP runtime per process: PyTorch ships its own copy of the runtime
        # (needed as libamdhip64.so); loading it after this library's (needed as
        # libamdhip64.so.7) would start a second runtime that finds no device.
        # Loaded first, torch's copy satisfies this library's dependency too.
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    st = ctypes.c_int
    L.rh_last_error.restype = ctypes.c_char_p
    L.rh_device_count.restype = ctypes.c_int32
    L.rh_model_create.argtypes = [ctypes.POINTER(_abi.ModelDesc), ctypes.c_int, _vp,
                                  ctypes.POINTER(_vp)]
    L.rh_model_destroy.argtypes = [_vp]
    L.rh_model_colloc.argtypes = [_vp, _d]
    L.rh_model_counts.argtypes = [_vp, _i64]
    L.rh_matrix_assemble.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
    L.rh_matrix_assemble_ex.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_abi.MatrixOpts),
                                        ctypes.POINTER(_vp)]
    L.rh_matrix_layout.argtypes = [_vp, _i64, _i64]
    L.rh_matrix_destroy.argtypes = [_vp]
    L.rh_matrix_export.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _d]
    L.rh_matrix_device_ptr.argtypes = [_vp, ctypes.POINTER(_vp), _i32]
    L.rh_matrix_format.argtypes = [_vp, _i32, _i64]
    L.rh_matrix_reassemble.argtypes = [_vp]
    L.rh_relax.argtypes = [_vp, _d, _d, _d, _d, ctypes.POINTER(_abi.RelaxOpts),
                           ctypes.POINTER(_abi.RelaxResult)]
    L.rh_relax_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_abi.RelaxOpts),
                               ctypes.POINTER(_abi.RelaxResult)]
    L.rh_relax_tab.argtypes = [_vp, _d, _d, ctypes.c_int64, _d, _d, _d, ctypes.POINTER(_abi.RelaxOpts),
                               ctypes.POINTER(_abi.RelaxResult)]
    L.rh_relax_tab_dev.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                   ctypes.POINTER(_abi.RelaxOpts), ctypes.POINTER(_abi.RelaxResult)]
    L.rh_field.argtypes = [_vp, _d, _d, ctypes.c_int64, _d, _d, ctypes.c_int]
    L.rh_field_dev.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp, ctypes.c_int]
    L.rh_field_dev_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp, ctypes.c_int]
    L.rh_current_field.argtypes = [ctypes.c_int, _vp, ctypes.c_int64, _i32, _d, _d, _d, _d, _d,
                                   ctypes.c_int64, _d]
    L.rh_current_field_dev.argtypes = [ctypes.c_int, _vp, ctypes.c_int64, _i32, _d, _d, _d, _d,
                                       _vp, ctypes.c_int64, _vp]
    L.rh_model_vertex_count.argtypes = [_vp, _i64]
    L.rh_model_set_field_form.argtypes = [_vp, ctypes.c_int32]
    L.rh_model_field_form.argtypes = [_vp, _i32]
    L.rh_model_mesh_counts.argtypes = [_vp, _i64]
    L.rh_field_adjoint.argtypes = [_vp, _d, _d, ctypes.c_int64, _d, ctypes.c_int32, _d, _d, _d]
    L.rh_field_adjoint_dev.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, ctypes.c_int32, _vp,
                                       _vp, _vp]
    L.rh_relax_adjoint_dev.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(_abi.RelaxOpts),
                                       ctypes.POINTER(_abi.RelaxResult)]
    L.rh_relax_adjoint.argtypes = [_vp, _d, _d, _d, ctypes.POINTER(_abi.RelaxOpts),
                                   ctypes.POINTER(_abi.RelaxResult)]
    L.rh_set_fallback.argtypes = [ctypes.c_int32]
    L.rh_get_fallback.restype = ctypes.c_int32
    L.rh_last_backend.argtypes = [ctypes.c_int32]
    L.rh_last_backend.restype = ctypes.c_int32
    L.rh_set_timing.argtypes = [ctypes.c_int32]
    L.rh_kernel_time.argtypes = [ctypes.c_char_p, _d, _i64]
    for name in ["rh_model_create", "rh_model_destroy", "rh_model_colloc", "rh_model_counts",
                 "rh_matrix_assemble", "rh_matrix_destroy", "rh_matrix_export",
                 "rh_matrix_device_ptr", "rh_matrix_reassemble", "rh_relax", "rh_relax_dev",
                 "rh_field", "rh_field_dev", "rh_field_dev_noalloc", "rh_current_field", "rh_current_field_dev",
                 "rh_set_timing", "rh_kernel_time", "rh_model_vertex_count",
                 "rh_field_adjoint", "rh_field_adjoint_dev", "rh_relax_adjoint_dev",
                 "rh_relax_adjoint", "rh_model_set_field_form", "rh_model_field_form",
                 "rh_model_mesh_counts", "rh_matrix_assemble_ex", "rh_matrix_layout",
                 "rh_relax_tab", "rh_relax_tab_dev", "rh_set_fallback", "rh_matrix_format"]:
        getattr(L, name).restype = st
    _RAW = L
    return L


def check(rc):
    if rc != RH_OK:
        raise RadHipError("radhip error %d: %s" % (rc, _load().rh_last_error().decode()))


def _p(a):
    return a.ctypes.data_as(_d)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def stream_handle():
    """The current torch stream when torch's GPU context is initialised, else
    the null stream."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None:
        try:
            if torch.cuda.is_initialized():
                return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:
            pass
    return ctypes.c_void_p(0)


class Model:
    """Device-resident magnetizable elements (rh_model)."""

    def __init__(self, packed, device=0, stream=None):
        L = lib()
        self.packed = packed
        self.n = len(packed["type"])
        desc, self._keep = _abi.model_desc(packed)
        self.h = _vp()
        self.device = device
        check(L.rh_model_create(ctypes.byref(desc), device,
                                stream if stream is not None else stream_handle(),
                                ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().rh_model_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def colloc(self):
        X = np.zeros((self.n, 3))
        check(lib().rh_model_colloc(self.h, _p(X)))
        return X

    def counts(self):
        c = np.zeros(8, np.int64)
        check(lib().rh_model_counts(self.h, c.ctypes.data_as(_i64)))
        return c

    def assemble(self, prec=RH_MIXED, **placement):
        return Matrix(self, prec, **placement)

    def set_field_form(self, form):
        """0: automatic (conforming-mesh form when held), 1: per element."""
        check(lib().rh_model_set_field_form(self.h, int(form)))

    def field_form(self):
        f = np.zeros(1, np.int32)
        check(lib().rh_model_field_form(self.h, f.ctypes.data_as(_i32)))
        return int(f[0])

    def mesh_counts(self):
        """(unique edges, unique faces, polyhedra) of the mesh form, zeros if none."""
        c = np.zeros(3, np.int64)
        check(lib().rh_model_mesh_counts(self.h, c.ctypes.data_as(_i64)))
        return tuple(int(v) for v in c)

    def vertex_count(self):
        nv = np.zeros(1, np.int64)
        check(lib().rh_model_vertex_count(self.h, nv.ctypes.data_as(_i64)))
        return int(nv[0])

    def field_adjoint(self, M, pts, gB, field="b", want_m=True, want_v=False, want_x=False):
        """Reverse-mode products of the field at pts for weights gB = dL/dB
        (field 'b') or dL/dH ('h'): returns (gM [n,3], gV [nvert,3], gX [np,3]),
        None where not requested."""
        pts = _c(pts).reshape(-1, 3)
        gB = _c(gB).reshape(-1, 3)
        Mv = _c(M).reshape(-1) if M is not None else None
        gM = np.zeros((self.n, 3)) if want_m else None
        gV = np.zeros((self.vertex_count(), 3)) if want_v else None
        gX = np.zeros_like(pts) if want_x else None
        check(lib().rh_field_adjoint(self.h, _p(Mv) if Mv is not None else None, _p(pts), len(pts),
                                     _p(gB), 0 if field == "b" else 1,
                                     _p(gM) if want_m else None, _p(gV) if want_v else None,
                                     _p(gX) if want_x else None))
        return gM, gV, gX

    def field(self, M, pts, want_b=True, want_h=True, precision=RH_FP64):
        M = _c(M).reshape(-1)
        pts = _c(pts).reshape(-1, 3)
        B = np.zeros_like(pts) if want_b else None
        H = np.zeros_like(pts) if (want_h or not want_b) else None
        check(lib().rh_field(self.h, _p(M), _p(pts), len(pts), _p(B) if B is not None else None,
                             _p(H) if H is not None else None, precision))
        return B, H


class Matrix:
    """Device-resident interaction matrix (rh_matrix)."""

    def __init__(self, model, prec=RH_MIXED, device_bytes=0, resident_fraction=-1.0,
                 allow_streaming=False, fp64_integrals=False, dense_rows=False):
        """device_bytes / resident_fraction / allow_streaming / fp64_integrals /
        dense_rows: rh_matrix_opts (rows beyond the device budget are streamed
        from pinned host memory when allowed; allow_streaming="policy" follows
        rh_get_fallback; fp64_integrals integrates every block of an RH_MIXED
        matrix in FP64 instead of the far ones in FP32; dense_rows stores
        row-major 3x3 blocks even where the symmetric-block layout applies)."""
        self.model = model
        self.n = model.n
        self.h = _vp()
        stream = _abi.RH_STREAM_BY_POLICY if allow_streaming == "policy" else (1 if allow_streaming else 0)
        o = _abi.MatrixOpts(int(device_bytes), float(resident_fraction), stream, 1 if fp64_integrals else 0,
                            1 if dense_rows else 0, 0)
        check(lib().rh_matrix_assemble_ex(model.h, int(prec), ctypes.byref(o), ctypes.byref(self.h)))

    def format(self):
        """(planes, ld): 0 for row-major 3x3 blocks (ld values per row), 5 or
        6 for the symmetric-block planes (ld values per plane)."""
        p = np.zeros(1, np.int32)
        ld = np.zeros(1, np.int64)
        check(lib().rh_matrix_format(self.h, p.ctypes.data_as(_i32), ld.ctypes.data_as(_i64)))
        return int(p[0]), int(ld[0])

    def layout(self):
        """(rows resident on the device, rows streamed from host memory)."""
        r = np.zeros(1, np.int64)
        s = np.zeros(1, np.int64)
        check(lib().rh_matrix_layout(self.h, r.ctypes.data_as(_i64), s.ctypes.data_as(_i64)))
        return int(r[0]), int(s[0])

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().rh_matrix_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reassemble(self):
        check(lib().rh_matrix_reassemble(self.h))

    def export(self, row0=0, nrows=None):
        n3 = 3 * self.n
        nrows = n3 - row0 if nrows is None else nrows
        out = np.zeros((nrows, n3))
        check(lib().rh_matrix_export(self.h, row0, nrows, _p(out)))
        return out

    def relax(self, mats, hext, M0, prec, max_iter, keep_m=False, omega0=0.0, refine=0, tab=None,
              method=9, krylov_m=0):
        """refine: FP64 refinement passes of a mixed-precision matrix
        (0 = library default of one pass, -1 = none); tab: node table of the
        tabulated laws that kind-3 records index (objects.pack_materials);
        method 9 (Anderson block Jacobi) or 11 (Newton-Krylov, GMRES restart
        krylov_m, 0 = 60)."""
        mats = _c(mats).reshape(-1)
        tab = _c(np.zeros(0) if tab is None else tab).reshape(-1)
        hext = _c(hext).reshape(-1)
        M = _c(M0).reshape(-1).copy()
        H = np.zeros_like(M)
        opts = _abi.RelaxOpts(float(prec), int(max_iter), 1 if keep_m else 0, float(omega0), int(method),
                              int(refine), int(krylov_m), 0)
        res = _abi.RelaxResult()
        check(lib().rh_relax_tab(self.h, _p(mats), _p(tab) if len(tab) else None, len(tab), _p(hext),
                                 _p(M), _p(H), ctypes.byref(opts), ctypes.byref(res)))
        return res, M.reshape(-1, 3), H.reshape(-1, 3)

    def relax_adjoint(self, mats, rhs, prec, max_iter, lam0=None):
        """Solves (I - Q^T K^T) lam = rhs (linear materials); returns (res, lam)."""
        mats = _c(mats).reshape(-1)
        rhs = _c(rhs).reshape(-1)
        lam = np.zeros_like(rhs) if lam0 is None else _c(lam0).reshape(-1).copy()
        opts = _abi.RelaxOpts(float(prec), int(max_iter), 0 if lam0 is None else 1, 0.0, 9, 0)
        res = _abi.RelaxResult()
        check(lib().rh_relax_adjoint(self.h, _p(mats), _p(rhs), _p(lam), ctypes.byref(opts),
                                     ctypes.byref(res)))
        return res, lam.reshape(-1, 3)


def current_field(cur, pts, device=0):
    """B of packed current sources (objects.pack_currents) at lab points."""
    pts = _c(pts).reshape(-1, 3)
    B = np.zeros_like(pts)
    ns = len(cur["kind"])
    if ns == 0:
        return B
    kind = np.ascontiguousarray(cur["kind"], dtype=np.int32)
    par = _c(cur["par"])
    tm, tv, ts = _c(cur["cM"]), _c(cur["cV"]), _c(cur["cS"])
    check(lib().rh_current_field(device, stream_handle(), ns, kind.ctypes.data_as(_i32), _p(par),
                                 _p(tm), _p(tv), _p(ts), _p(pts), len(pts), _p(B)))
    return B


def mat_M(rec, H, tab=None):
    """M(H) of a material record (radmater.h:224, :351, :644-657; tabulated
    laws radmater.cpp:619-646 with the node table tab)."""
    rec = np.asarray(rec, dtype=np.float64)
    H = np.asarray(H, dtype=np.float64)
    kind = int(rec[0])
    mr = rec[13:16]
    if kind == 0:
        return rec[1] * H + mr
    if kind == 1:
        return rec[4:13].reshape(3, 3) @ H + mr
    h = float(np.linalg.norm(H))
    if kind == 3:
        from .objects import tab_eval
        t = np.asarray(tab, dtype=np.float64)[3 * int(rec[1]):3 * int(rec[1] + rec[2])].reshape(-1, 3)
        return (tab_eval(t[:, 0], t[:, 1], t[:, 2], h) / h) * H + mr if h != 0 else mr.copy()
    f = 0.0
    for t in range(int(rec[1])):
        ms, ks = rec[2 + t], rec[5 + t]
        if ms != 0:
            f += ms * np.tanh(ks * h / ms)
    return (f / h) * H + mr if h != 0 else mr.copy()


FALLBACK_NAMES = ("cpu", "gpu_streaming", "break")


def set_fallback(mode):
    """Process-wide fallback policy of the library (rh_set_fallback): 0 'cpu',
    1 'gpu_streaming', 2 'break'."""
    check(_load().rh_set_fallback(int(mode)))


def get_fallback():
    return int(_load().rh_get_fallback())


def last_backend(path):
    """rh_last_backend: path 0 assembly, 1 field; -1 none, 1 device, 2 device
    with streamed rows."""
    return int(_load().rh_last_backend(int(path)))


def set_timing(on):
    check(lib().rh_set_timing(1 if on else 0))


def kernel_time(family):
    avg = ctypes.c_double()
    cnt = ctypes.c_int64()
    check(lib().rh_kernel_time(family.encode(), ctypes.byref(avg), ctypes.byref(cnt)))
    return avg.value, cnt.value
