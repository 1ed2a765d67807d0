"""ctypes front end of the CPU oracle (libradora.so, built from radora.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of the benchmark as the checker.  The product never imports this module.

The model argument of the model-level functions is a plain dict of numpy
arrays describing relaxable elements as the user built them (see
``prepare_model``); the oracle derives face frames, normals and centroids
itself (restating radvlpgn.cpp), independently of the product's packing.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_ci = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "libradora.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libradora.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ora_set_randomization.argtypes = [ctypes.c_int]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_recmag_Q.argtypes = [_d, _d, _d, _d]
        L.ora_recmag_BH.argtypes = [_d, _d, _d, _d, _d, _d]
        L.ora_reccur_BH.argtypes = [_d, _d, _d, _d, _d, _d]
        L.ora_polygon_S.argtypes = [_d, ctypes.c_int, ctypes.c_double, _d, _d, _d, _ci]
        L.ora_poly_prepare.argtypes = [_d, ctypes.c_int, _ci, _ci, ctypes.c_int, _ci, _d, _d,
                                       _d, _d, _d, _d]
        L.ora_poly_prepare.restype = ctypes.c_int
        L.ora_poly_Q.argtypes = [_d, ctypes.c_int, _ci, _d, _d, _d, _d, _d, _d]
        L.ora_poly_BH.argtypes = [_d, ctypes.c_int, _ci, _d, _d, _d, _d, _d, _d, _d, _d]
        margs = [ctypes.c_int64, _i32, _d, _d, _i64, _i32, _i64, _d, _d, _d, _d, _d, _d,
                 _i64, _d, _d, _d, _d]
        L.ora_model_assemble.argtypes = margs + [ctypes.c_int64, ctypes.c_int64, _d]
        L.ora_model_field.argtypes = margs + [_d, _d, ctypes.c_int64, _d, _d]
        L.ora_model_colloc.argtypes = margs + [_d]
        L.ora_mat_M.argtypes = [_d, _d, _d]
        L.ora_relax_m4.argtypes = [ctypes.c_int64, _f, _d, _d, _d, _d, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, _d]
        L.ora_relax_m4.restype = ctypes.c_int
        L.ora_arccur_BH.argtypes = [_d, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, _d, _d]
        L.ora_inv3.argtypes = [_d, _d]
        L.ora_set_tables.argtypes = [_d]
        L.ora_tab_prepare.argtypes = [_d, ctypes.c_int, _d, _d, _d]
        L.ora_tab_prepare.restype = ctypes.c_int
        L.ora_tab_eval.argtypes = [_d, ctypes.c_int, ctypes.c_double]
        L.ora_tab_eval.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _v(x, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


def set_randomization(on):
    lib().ora_set_randomization(1 if on else 0)


def set_threads(n):
    """Threads of the oracle's OpenMP loops (1 = the reference's serial loops)."""
    lib().ora_set_threads(int(n))


# ---------------------------------------------------------------- elements
def recmag_Q(P, C, D):
    Q = np.zeros(9)
    lib().ora_recmag_Q(_p(_v(P)), _p(_v(C)), _p(_v(D)), _p(Q))
    return Q.reshape(3, 3)


def recmag_BH(P, C, D, M):
    B, H = np.zeros(3), np.zeros(3)
    lib().ora_recmag_BH(_p(_v(P)), _p(_v(C)), _p(_v(D)), _p(_v(M)), _p(B), _p(H))
    return B, H


def reccur_BH(P, C, D, J):
    B, H = np.zeros(3), np.zeros(3)
    lib().ora_reccur_BH(_p(_v(P)), _p(_v(C)), _p(_v(D)), _p(_v(J)), _p(B), _p(H))
    return B, H


def arccur_BH(P, C, rmin, rmax, phimin, phimax, h, j, nseg):
    B, H = np.zeros(3), np.zeros(3)
    lib().ora_arccur_BH(_p(_v(P)), _p(_v(C)), rmin, rmax, phimin, phimax, h, j, int(nseg),
                        _p(B), _p(H))
    return B, H


class Poly:
    """A polyhedron prepared by the oracle (faces re-oriented, frames, centroid)."""

    def __init__(self, verts, faces):
        verts = _v(verts).reshape(-1, 3)
        face_len = np.array([len(f) for f in faces], dtype=np.intc)
        face_idx = np.array([i for f in faces for i in f], dtype=np.intc)
        nf = len(faces)
        self.nf = nf
        self.face_len = face_len
        self.face_idx = np.zeros_like(face_idx)
        self.R = np.zeros(9 * nf)
        self.Rinv = np.zeros(9 * nf)
        self.v2d = np.zeros(2 * len(face_idx))
        self.coordz = np.zeros(nf)
        self.c2d = np.zeros(2 * nf)
        self.centroid = np.zeros(3)
        rc = lib().ora_poly_prepare(_p(verts), len(verts), _p(face_len, _ci), _p(face_idx, _ci),
                                    nf, _p(self.face_idx, _ci), _p(self.R), _p(self.Rinv),
                                    _p(self.v2d), _p(self.coordz), _p(self.c2d),
                                    _p(self.centroid))
        if rc != 0:
            raise ValueError("degenerate polyhedron face (code %d)" % rc)

    def Q(self, P):
        Q = np.zeros(9)
        lib().ora_poly_Q(_p(_v(P)), self.nf, _p(self.face_len, _ci), _p(self.R), _p(self.Rinv),
                         _p(self.v2d), _p(self.coordz), _p(self.c2d), _p(Q))
        return Q.reshape(3, 3)

    def BH(self, P, M):
        B, H = np.zeros(3), np.zeros(3)
        lib().ora_poly_BH(_p(_v(P)), self.nf, _p(self.face_len, _ci), _p(self.R),
                          _p(self.Rinv), _p(self.v2d), _p(self.coordz), _p(self.c2d),
                          _p(_v(M)), _p(B), _p(H))
        return B, H


# ---------------------------------------------------------------- models
def prepare_model(raw):
    """Turn a raw element description into the oracle's flat model.

    raw keys: type[n] (0 RecMag, 1 polyhedron), rec_c[n,3], rec_d[n,3],
    vert_off[n+1], verts[V,3], face_off[n+1], face_len[F], face_idx[sum len]
    (0-based, per element), copy_off[n+1], cM[C,3,3], cV[C,3], cS[C].
    """
    n = len(raw["type"])
    typ = _v(raw["type"], np.int32)
    rec_c = _v(raw["rec_c"]).reshape(n, 3)
    rec_d = _v(raw["rec_d"]).reshape(n, 3)
    pf_off = [0]
    fl_all, R_all, Ri_all, v2d_all, cz_all, c2d_all = [], [], [], [], [], []
    fvo = [0]
    centroid = np.zeros((n, 3))
    vo, fo = raw["vert_off"], raw["face_off"]
    fl = raw["face_len"]
    fidx = raw["face_idx"]
    # face_idx offsets per face
    fstart = np.concatenate([[0], np.cumsum(fl)])
    for e in range(n):
        if typ[e] == 0:
            centroid[e] = rec_c[e]
            pf_off.append(pf_off[-1])
            continue
        verts = raw["verts"][vo[e]:vo[e + 1]]
        faces = [list(fidx[fstart[f]:fstart[f + 1]]) for f in range(fo[e], fo[e + 1])]
        p = Poly(verts, faces)
        centroid[e] = p.centroid
        fl_all.append(p.face_len)
        R_all.append(p.R)
        Ri_all.append(p.Rinv)
        v2d_all.append(p.v2d)
        cz_all.append(p.coordz)
        c2d_all.append(p.c2d)
        for L in p.face_len:
            fvo.append(fvo[-1] + int(L))
        pf_off.append(pf_off[-1] + p.nf)
    cat = lambda lst, dt=np.float64: (np.concatenate(lst).astype(dt) if lst
                                      else np.zeros(1, dt))
    cM = _v(raw["cM"]).reshape(-1, 3, 3)
    cMinv = np.stack([np.linalg.inv(m) for m in cM]) if len(cM) else np.zeros((0, 3, 3))
    # the oracle inverts copy matrices with Radia's cofactor formula
    cMinv = np.zeros_like(cM)
    for c in range(len(cM)):
        o = np.zeros(9)
        lib().ora_inv3(_p(_v(cM[c]).ravel()), _p(o))
        cMinv[c] = o.reshape(3, 3)
    return dict(
        n=n, type=typ, rec_c=_v(rec_c), rec_d=_v(rec_d),
        poly_face_off=_v(pf_off, np.int64), face_len=cat(fl_all, np.int32),
        face_vtx_off=_v(fvo, np.int64), R=cat(R_all), Rinv=cat(Ri_all), v2d=cat(v2d_all),
        coordz=cat(cz_all), c2d=cat(c2d_all), centroid=_v(centroid),
        copy_off=_v(raw["copy_off"], np.int64), cM=_v(cM), cMinv=_v(cMinv),
        cV=_v(raw["cV"]).reshape(-1, 3), cS=_v(raw["cS"]))


def _margs(m):
    return [m["n"], _p(m["type"], _i32), _p(m["rec_c"]), _p(m["rec_d"]),
            _p(m["poly_face_off"], _i64), _p(m["face_len"], _i32),
            _p(m["face_vtx_off"], _i64), _p(m["R"]), _p(m["Rinv"]), _p(m["v2d"]),
            _p(m["coordz"]), _p(m["c2d"]), _p(m["centroid"]), _p(m["copy_off"], _i64),
            _p(m["cM"]), _p(m["cMinv"]), _p(m["cV"]), _p(m["cS"])]


def assemble(m, rows=None):
    """Interaction matrix rows [i0, i1) in FP64, shape (3*(i1-i0), 3n)."""
    n = m["n"]
    i0, i1 = (0, n) if rows is None else rows
    Q = np.zeros((3 * (i1 - i0), 3 * n))
    lib().ora_model_assemble(*_margs(m), i0, i1, _p(Q))
    return Q


def colloc(m):
    X = np.zeros((m["n"], 3))
    lib().ora_model_colloc(*_margs(m), _p(X))
    return X


def field(m, Mloc, pts):
    pts = _v(pts).reshape(-1, 3)
    B = np.zeros_like(pts)
    H = np.zeros_like(pts)
    lib().ora_model_field(*_margs(m), _p(_v(Mloc)), _p(pts), len(pts), _p(B), _p(H))
    return B, H


# ---------------------------------------------------------------- materials / relaxation
MAT_STRIDE = 16


_TABLES = [None]


def tab_prepare(points):
    """Node table (x, y, d) of a tabulated law from (H, M) points (radmater.cpp:460-609)."""
    pts = _v(points).reshape(-1, 2)
    n = len(pts)
    x, y, d = np.zeros(n + 1), np.zeros(n + 1), np.zeros(n + 1)
    L = lib().ora_tab_prepare(_p(pts.ravel()), n, _p(x), _p(y), _p(d))
    if L < 0:
        raise ValueError("a tabulated law needs at least 4 nodes")
    return x[:L], y[:L], d[:L]


def tab_eval(table, h):
    """|M|(|H|) of a node table given as flat (x, y, d) triples."""
    t = _v(table).ravel()
    return lib().ora_tab_eval(_p(t), len(t) // 3, float(h))


def set_tables(tab):
    """Flat (x, y, d) triples that kind-3 material records index."""
    t = None if tab is None else _v(tab).ravel().copy()
    _TABLES[0] = t  # keep alive while the oracle holds the pointer
    lib().ora_set_tables(_p(t) if t is not None else None)


def mat_M(mat, H):
    M = np.zeros(3)
    lib().ora_mat_M(_p(_v(mat)), _p(_v(H)), _p(M))
    return M


def relax_m4(Qf, mats, Hext, prec, max_iter, M0=None):
    """Method-4 Gauss-Seidel relaxation on a float matrix; returns (M, H, iters, stats)."""
    Qf = np.ascontiguousarray(Qf, dtype=np.float32)
    n = Qf.shape[0] // 3
    M = np.zeros(3 * n) if M0 is None else _v(M0).ravel().copy()
    H = np.zeros(3 * n)
    stats = np.zeros(3)
    it = lib().ora_relax_m4(n, _p(Qf, _f), _p(_v(mats).ravel()), _p(_v(Hext).ravel()),
                            _p(M), _p(H), prec, int(max_iter), 0 if M0 is None else 1,
                            _p(stats))
    return M.reshape(n, 3), H.reshape(n, 3), it, stats
