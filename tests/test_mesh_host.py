"""CPU checks of the conforming-mesh field form (rh_mesh.hpp, host build):
the regrouped sum over unique edges and faces equals the per-element sum, and
the unique-entity counts are those of the mesh."""
import ctypes
import os

import numpy as np
import pytest

from radiacuda_amd import _abi
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from radiacuda_amd import workloads as W
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


def _fields(L, P, M, pts, fp32=0):
    desc, keep = _abi.model_desc(P)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    out = {}
    for name in ("elem", "mesh"):
        B = np.zeros_like(pts)
        H = np.zeros_like(pts)
        if name == "elem":
            rc = L.rhc_field(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                             ctypes.c_int64(len(pts)), ctypes.c_int32(fp32), B.ctypes.data_as(dp),
                             H.ctypes.data_as(dp))
            counts = None
        else:
            counts = np.zeros(3, np.int64)
            rc = L.rhc_field_mesh(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                                  ctypes.c_int64(len(pts)), ctypes.c_int32(fp32), B.ctypes.data_as(dp),
                                  H.ctypes.data_as(dp), counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        assert rc == 0, rc
        out[name] = (B, H, counts)
    return out


def test_mesh_form_tets_and_blocks_equals_per_element(L):
    rng = np.random.default_rng(31)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=32)
    ids.append(rad.ObjRecMag([20, 3, 2], [3, 4, 5], list(rng.uniform(-1, 1, 3))))
    P, items = models.packed(rad.ObjCnt(ids))
    M = rng.uniform(-1, 1, (len(items), 3))
    # points inside and outside the mesh (none on a face plane)
    pts = np.concatenate([rng.uniform(0.2, 7.8, (80, 3)), rng.uniform(-10, 25, (80, 3))])
    f = _fields(L, P, M, pts)
    Be, He, _ = f["elem"]
    Bm, Hm, counts = f["mesh"]
    assert counts[2] == len(items) - 1
    assert counts[1] < 4 * counts[2]  # interior faces are shared
    assert np.abs(Hm - He).max() <= 1e-12 * np.abs(He).max()
    assert np.abs(Bm - Be).max() <= 1e-12 * np.abs(Be).max()


def test_mesh_form_hexahedra_equals_per_element(L):
    rng = np.random.default_rng(33)
    top, _pts = W.halbach(n_seg=4, n_az=2, n_rad=2, n_ax=2, n_points=10)
    P, items = models.packed(top)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = np.concatenate([rng.uniform(-45, 45, (100, 3)),
                          np.stack([31 * np.cos(np.linspace(0.1, 6.1, 40)),
                                    31 * np.sin(np.linspace(0.1, 6.1, 40)),
                                    np.linspace(-20, 20, 40)], 1)])
    for fp32, tol in ((0, 1e-12), (1, 1e-5)):
        f = _fields(L, P, M, pts, fp32)
        Be, He, _ = f["elem"]
        Bm, Hm, counts = f["mesh"]
        assert np.abs(Hm - He).max() <= tol * np.abs(He).max()
        assert np.abs(Bm - Be).max() <= tol * np.abs(Be).max()
    assert counts[1] < 6 * counts[2]


def test_mesh_counts_c_yoke(L):
    top, iron, coil, info = W.c_yoke()
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    M = np.tile([0.0, 0.0, 1.0], (len(items), 3 // 3))
    f = _fields(L, P, M, W.gap_grid(info, 2))
    _B, _H, counts = f["mesh"]
    assert list(counts) == [11454, 17200, 8000]


def test_mesh_form_with_transform_copies_equals_per_element(L):
    """Every transform copy is a mesh element in the laboratory frame: a tet
    mesh under a field-inverting mirror (TrfZerPara), a plain mirror
    (TrfZerPerp) and a 4-fold rotation (TrfMlt), plus a block, give the
    per-element sum.  Faces on the mirror planes are shared between an
    element and its image."""
    rng = np.random.default_rng(35)
    rad.UtiDelAll()
    ids = models.tet_mesh((2, 2, 2), 3.0, origin=(0.0, 0.0, 0.0), jitter=0.1, seed=36)
    mesh = rad.ObjCnt(ids)
    rad.TrfZerPara(mesh, [0, 0, 0], [0, 0, 1])
    rad.TrfZerPerp(mesh, [0, 0, 0], [1, 0, 0])
    blk = rad.ObjRecMag([15, 2, 1], [3, 4, 5], [0.1, 0.2, 0.3])
    rot = rad.ObjCnt([rad.ObjPolyhdr([[20, 0, 0], [24, 0, 0], [20, 4, 0], [20, 0, 4]],
                                     [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]], [0.2, -0.3, 0.5])])
    rad.TrfMlt(rot, rad.TrfRot([0, 0, 0], [0, 0, 1], np.pi / 2), 4)
    top = rad.ObjCnt([mesh, blk, rot])
    P, items = models.packed(top)
    M = rng.uniform(-1, 1, (len(items), 3))
    pts = np.concatenate([rng.uniform(-5.8, 5.8, (80, 3)), rng.uniform(-30, 30, (120, 3))])
    for fp32, tol in ((0, 1e-12), (1, 1e-5)):
        f = _fields(L, P, M, pts, fp32)
        Be, He, _ = f["elem"]
        Bm, Hm, counts = f["mesh"]
        assert np.abs(Hm - He).max() <= tol * np.abs(He).max(), fp32
        assert np.abs(Bm - Be).max() <= tol * np.abs(Be).max(), fp32
    # the mesh's tetrahedra x 4 copies + 4 rotated tetrahedra (the block is not a mesh element)
    assert counts[2] == len(ids) * 4 + 4
    # faces on the two mirror planes are shared: fewer unique faces than without sharing
    assert counts[1] < 4 * counts[2] * 0.6
    rad.UtiDelAll()


def _mesh_stats(L, top):
    items = [it for it in O.walk(rad._reg, top)
             if isinstance(it[0], (O.RecMag, O.Polyhedron)) and not getattr(it[0], "is_current", False)]
    desc, keep = _abi.model_desc(O.pack_elements(items))
    out = np.zeros(7, np.int64)
    assert L.rhc_mesh_stats(ctypes.byref(desc), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))) == 0
    return dict(zip(("edges", "free", "faces", "tri", "tri_cont", "quad", "quad_cont"), out.tolist()))


def test_sweep_order_reuses_offsets_on_the_bench_meshes(L):
    """The packer's strips and side ownership: on the C-yoke tetrahedra every
    edge is evaluated by a triangle and most triangles continue a strip; on the
    hexahedral Halbach ring almost every edge is owned by a quadrilateral and
    most quadrilaterals continue a strip (each continued face computes one or
    two vertex offsets instead of three or four)."""
    rad.UtiDelAll()
    _top, iron, _coil, _info = W.c_yoke()
    s = _mesh_stats(L, iron)
    assert s["quad"] == 0 and s["tri"] == s["faces"] == 17200 and s["edges"] == 11454
    assert s["free"] == 0 and s["tri_cont"] >= 0.85 * s["tri"]
    rad.UtiDelAll()
    top, _pts = W.halbach(n_points=10)
    s = _mesh_stats(L, top)
    assert s["tri"] == 0 and s["quad"] == s["faces"]
    assert s["free"] <= 0.01 * s["edges"] and s["quad_cont"] >= 0.7 * s["quad"]
    rad.UtiDelAll()
