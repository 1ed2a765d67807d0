"""CPU checks of the adjoint math (rh_adjoint.hpp, host build in _rhcheck.so).

* The triangulated surface field the vertex derivatives are taken of equals
  the face-form field of the forward kernels for planar faces.
* Vertex and point gradients from the dual-number derivatives equal central
  finite differences of the oracle's field (radora.c, the reference's face
  formula) for tetrahedra, where every face stays planar under any vertex
  displacement, and for hexahedra along displacements that keep the faces
  planar (rigid translation, scaling about the centroid).
* For arbitrary displacements of hexahedron vertices the gradient equals
  finite differences of the forward path itself, which evaluates faces that
  are no longer planar as their fan of triangles.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _abi
from radiacuda_amd import rad
from tests import models

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "radiacuda_amd", "_rhcheck.so")
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(LIB):
        from radiacuda_amd import build
        build.build()
    return ctypes.CDLL(LIB)


@pytest.fixture(autouse=True)
def _rand_off():
    OR.set_randomization(False)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _adj_field(L, P, M, pts):
    desc, keep = _abi.model_desc(P)
    pts, M = _c(pts), _c(M)
    H = np.zeros_like(pts)
    assert L.rhc_adj_field(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                           ctypes.c_int64(len(pts)), H.ctypes.data_as(dp)) == 0
    return H


def _adj_grad(L, P, M, pts, g):
    desc, keep = _abi.model_desc(P)
    pts, M, g = _c(pts), _c(M), _c(g)
    gV = np.zeros((len(P["verts"]), 3))
    gX = np.zeros_like(pts)
    assert L.rhc_adj_grad(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                          ctypes.c_int64(len(pts)), g.ctypes.data_as(dp), gV.ctypes.data_as(dp),
                          gX.ctypes.data_as(dp)) == 0
    return gV, gX


def _fwd_H(L, P, M, pts):
    desc, keep = _abi.model_desc(P)
    pts, M = _c(pts), _c(M)
    B = np.zeros_like(pts)
    H = np.zeros_like(pts)
    assert L.rhc_field(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                       ctypes.c_int64(len(pts)), ctypes.c_int32(0), B.ctypes.data_as(dp),
                       H.ctypes.data_as(dp)) == 0
    return H


def _oracle_H(P, M, pts):
    _B, H = OR.field(OR.prepare_model(P), M, pts)
    return H


def _with_verts(P, V):
    Q = dict(P)
    Q["verts"] = V
    return Q


def _hex_model(rng, count=3):
    ids = []
    for k in range(count):
        q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        rot = q * np.sign(np.linalg.det(q))
        ids.append(models.hex_block([6.0 * k, 1.0, -2.0], [3.0, 2.0, 4.0], rng.uniform(-1, 1, 3), rot))
    return rad.ObjCnt(ids)


def test_triangulated_field_equals_face_form(L):
    rng = np.random.default_rng(3)
    top = rad.ObjCnt([rad.ObjCnt(models.tet_mesh((2, 1, 1), 4.0, jitter=0.1, seed=4)), _hex_model(rng)])
    P, items = models.packed(top)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = rng.uniform(-6, 16, (60, 3))
    Ht = _adj_field(L, P, M, pts)
    Hf = _fwd_H(L, P, M, pts)
    assert np.abs(Ht - Hf).max() <= 1e-12 * np.abs(Hf).max()


def _fd_check(L, P, M, pts, g, ref_H, coords, h, tol):
    gV, _gX = _adj_grad(L, P, M, pts, g)
    V0 = _c(P["verts"])
    worst = 0.0
    for (v, c) in coords:
        Vp, Vm = V0.copy(), V0.copy()
        Vp[v, c] += h
        Vm[v, c] -= h
        fd = (np.sum(g * ref_H(_with_verts(P, Vp), M, pts)) -
              np.sum(g * ref_H(_with_verts(P, Vm), M, pts))) / (2 * h)
        scale = max(np.abs(gV).max(), 1e-300)
        worst = max(worst, abs(fd - gV[v, c]) / scale)
    assert worst <= tol, worst
    return gV


def _rotated(P, rng):
    """The model turned by a random rotation.  Faces whose outward normal lies
    close to -z are avoided that way: the reference's face rotation divides by
    1 + n_z (radvlpgn.cpp:357-371, restated in the oracle) and loses accuracy
    there, which finite differences of tilted axis-aligned faces expose."""
    q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.linalg.det(q))
    return _with_verts(P, _c(P["verts"]) @ q.T)


def test_vertex_gradient_tets_vs_oracle_fd(L):
    rng = np.random.default_rng(11)
    top = rad.ObjCnt(models.tet_mesh((1, 1, 1), 4.0, jitter=0.1, seed=12))
    P, items = models.packed(top)
    P = _rotated(P, rng)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = rng.uniform(-3, 7, (25, 3))
    g = rng.normal(size=pts.shape)
    coords = [(v, c) for v in range(len(P["verts"])) for c in range(3)]
    _fd_check(L, P, M, pts, g, _oracle_H, coords, 1e-5, 1e-7)


def test_point_gradient_vs_oracle_fd(L):
    rng = np.random.default_rng(13)
    top = rad.ObjCnt([rad.ObjCnt(models.tet_mesh((1, 1, 1), 4.0, jitter=0.1, seed=14)), _hex_model(rng, 2)])
    P, items = models.packed(top)
    P = _rotated(P, rng)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = rng.uniform(-3, 9, (12, 3))
    g = rng.normal(size=pts.shape)
    _gV, gX = _adj_grad(L, P, M, pts, g)
    h = 1e-5
    for p in range(len(pts)):
        for c in range(3):
            xp, xm = pts.copy(), pts.copy()
            xp[p, c] += h
            xm[p, c] -= h
            fd = (np.sum(g * _oracle_H(P, M, xp)) - np.sum(g * _oracle_H(P, M, xm))) / (2 * h)
            assert abs(fd - gX[p, c]) <= 1e-7 * np.abs(gX).max()


def test_vertex_gradient_hex_planar_motions_vs_oracle_fd(L):
    """Rigid translation and scaling of one hexahedron keep its faces planar,
    so the oracle's face formula pins the directional derivatives."""
    rng = np.random.default_rng(17)
    top = _hex_model(rng, 3)
    P, items = models.packed(top)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = rng.uniform(-4, 16, (30, 3))
    g = rng.normal(size=pts.shape)
    gV, _gX = _adj_grad(L, P, M, pts, g)
    V0 = _c(P["verts"])
    vo = P["vert_off"]
    h = 1e-5
    for e in range(len(items)):
        sl = slice(vo[e], vo[e + 1])
        cen = V0[sl].mean(axis=0)
        for d in [np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]), None]:
            dirs = np.zeros_like(V0)
            dirs[sl] = (V0[sl] - cen) if d is None else d
            fd = (np.sum(g * _oracle_H(_with_verts(P, V0 + h * dirs), M, pts)) -
                  np.sum(g * _oracle_H(_with_verts(P, V0 - h * dirs), M, pts))) / (2 * h)
            an = np.sum(gV * dirs)
            assert abs(fd - an) <= 1e-7 * max(abs(an), np.abs(gV).max()), (e, d, fd, an)


def test_vertex_gradient_hex_any_motion_vs_forward_fd(L):
    """Single-vertex displacements make faces non-planar; the forward path then
    evaluates them as fan triangles, and its finite differences match."""
    rng = np.random.default_rng(19)
    top = _hex_model(rng, 2)
    P, items = models.packed(top)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = rng.uniform(-4, 10, (20, 3))
    g = rng.normal(size=pts.shape)
    coords = [(v, c) for v in range(len(P["verts"])) for c in range(3)]
    _fd_check(L, P, M, pts, g, lambda Q, MM, X: _fwd_H(L, Q, MM, X), coords, 1e-6, 1e-6)
