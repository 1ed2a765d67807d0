"""CPU build of the exact kernel math (rh_math.hpp / rh_eval.hpp / rh_current.hpp,
compiled for the host into _rhcheck.so) against the oracle: validates the
formulation and the packing without a GPU."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _abi
from radiacuda_amd import objects as O
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


def _assemble(L, P, fp32=0):
    desc, keep = _abi.model_desc(P)
    n = len(P["type"])
    Q = np.zeros((3 * n, 3 * n))
    rc = L.rhc_assemble(ctypes.byref(desc), ctypes.c_int64(0), ctypes.c_int64(n), ctypes.c_int32(fp32),
                        Q.ctypes.data_as(dp))
    assert rc == 0
    return Q


def _field(L, P, M, pts, fp32=0, kinds=False):
    desc, keep = _abi.model_desc(P)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    B = np.zeros_like(pts)
    H = np.zeros_like(pts)
    fn = L.rhc_field_kinds if kinds else L.rhc_field
    rc = fn(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
            ctypes.c_int64(len(pts)), ctypes.c_int32(fp32), B.ctypes.data_as(dp),
            H.ctypes.data_as(dp))
    assert rc == 0
    return B, H


def _block_rel(A, B, n):
    d = np.abs(A - B).reshape(n, 3, n, 3).max(axis=(1, 3))
    s = np.abs(B).reshape(n, 3, n, 3).max(axis=(1, 3))
    return (d / np.maximum(s, 1e-300)).max()


def test_assembly_formulation_matches_oracle(L):
    top = models.mixed_model(seed=3)
    P, _ = models.packed(top)
    n = len(P["type"])
    Q = _assemble(L, P)
    Qo, jit = models.oracle_assemble(P)
    clean, jittered = models.block_rel(Q, Qo, n, jit)
    assert np.isfinite(Q).all()
    assert clean <= 1e-10
    assert jittered <= 1e-6


def test_field_formulation_matches_oracle(L):
    top = models.mixed_model(seed=4)
    P, _ = models.packed(top)
    rng = np.random.default_rng(9)
    pts = rng.uniform(-12, 16, (300, 3))
    Bo, Ho, _bad = models.oracle_field(P, P["M"], pts)
    for kinds in (False, True):  # generic dispatch, and the per-kind path of the kernels
        B, H = _field(L, P, P["M"], pts, kinds=kinds)
        assert np.abs(B - Bo).max() <= 1e-11 * np.abs(Bo).max()
        assert np.abs(H - Ho).max() <= 1e-11 * np.abs(Ho).max()


def test_fp32_field_error_is_bounded(L):
    """FP32 evaluation of the integrals (the mixed path): stated error vs FP64."""
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=5)
    top = rad.ObjCnt(ids)
    P, _ = models.packed(top)
    rng = np.random.default_rng(2)
    pts = rng.uniform(-4, 16, (200, 3))
    B64, _ = _field(L, P, P["M"], pts, 0)
    B32, _ = _field(L, P, P["M"], pts, 1, kinds=True)
    err = np.abs(B32 - B64).max() / np.abs(B64).max()
    # the reference bounds its FP32 face kernel at 5e-4 (tests/test_gpu_fld_fp32.py:72);
    # the cancellation-free FP32 edge form stays well inside 1e-6 here
    assert err < 1e-6


def test_current_sources_match_oracle(L):
    rot = O.Trf.rotation([1, 1, 1], [0.2, 0.3, 1], 0.4)
    cur = [(O.RecMag([1, 2, 3], [2, 3, 4], (0, 0, 0), (1.5, -2, 0.7)), 0, ()),
           (O.ArcCur([0, 0, 1], 5, 8, 0.2, 1.7, 3, 10, 2.5), 1, ([(rot, 1)],))]
    pc = O.pack_currents(cur)
    ns = len(pc["kind"])
    cp = np.zeros((ns, 24))
    for s in range(ns):
        cp[s, :9] = np.linalg.inv(pc["cM"][s]).ravel()
        cp[s, 9:12] = pc["cV"][s]
        cp[s, 12:21] = (pc["cS"][s] * pc["cM"][s]).ravel()
    rng = np.random.default_rng(4)
    pts = np.ascontiguousarray(rng.uniform(-10, 10, (100, 3)))
    B = np.zeros_like(pts)
    kind = np.ascontiguousarray(pc["kind"])
    par = np.ascontiguousarray(pc["par"])
    L.rhc_current(ctypes.c_int64(ns), kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                  par.ctypes.data_as(dp), cp.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                  ctypes.c_int64(len(pts)), B.ctypes.data_as(dp))
    Bo = np.zeros_like(pts)
    for i, p in enumerate(pts):
        b, _ = OR.reccur_BH(p, [1, 2, 3], [2, 3, 4], (1.5, -2, 0.7))
        Bo[i] += b
        xl = np.linalg.solve(rot.M, p - rot.V)
        b, _ = OR.arccur_BH(xl, [0, 0, 1], 5, 8, 0.2, 1.7, 3, 2.5, 10)
        Bo[i] += rot.M @ b
    assert np.abs(B - Bo).max() <= 1e-12 * np.abs(Bo).max()


def test_symmetric_model_equals_explicit_mirror(L):
    """A TrfZerPara symmetry equals adding the mirrored elements explicitly with
    inverted, mirrored magnetization (radentry.cpp:1802-1840)."""
    rng = np.random.default_rng(8)
    ids = models.tet_mesh((1, 1, 1), 3.0, origin=(1, 1, 2), jitter=0.1, seed=8)
    a = rad.ObjCnt(ids)
    rad.TrfZerPara(a, [0, 0, 0], [0, 0, 1])
    P, items = models.packed(a)
    rad.UtiDelAll()
    ids2 = models.tet_mesh((1, 1, 1), 3.0, origin=(1, 1, 2), jitter=0.1, seed=8)
    mir = []
    S = np.diag([1.0, 1.0, -1.0])
    for k in ids2:
        o = rad._reg[k]
        v = o.verts @ S.T
        mir.append(rad.ObjPolyhdr(v.tolist(), [[i + 1 for i in f] for f in o.faces],
                                  list(-(S @ o.M))))
    b = rad.ObjCnt(ids2 + mir)
    P2, _ = models.packed(b)
    pts = rng.uniform(-6, 6, (100, 3))
    B1, _ = _field(L, P, P["M"], pts)
    B2, _ = _field(L, P2, P2["M"], pts)
    assert np.abs(B1 - B2).max() <= 1e-12 * np.abs(B2).max()


def test_kernel_math_matches_50_digit_values_where_reference_form_loses_digits(L):
    """Tetrahedra of the jittered C-yoke whose faces have edges nearly parallel
    to a face-frame axis, and tetrahedra with a face whose outward normal is
    close to -z: the kernel formulation matches 50-digit evaluations
    (tests/golden/mp50_tets.json) to <= 1e-11 relative.  The reference's
    forms, restated by the oracle, lose accuracy on both kinds: the
    y = kx + b edge form (radplnr2.cpp:139-170) up to ~1e-3 of these far
    elements' contributions, and the face rotation (radvlpgn.cpp:357-371),
    which divides by 1 + n_z, about 4e-4 on the -z cases."""
    import json
    gold = json.load(open(os.path.join(HERE, "golden", "mp50_tets.json")))
    worst_k, worst_o, worst_oz = 0.0, 0.0, 0.0
    for c in gold["cases"]:
        o = O.Polyhedron(c["verts"], c["faces"], c["M"])
        P = O.pack_elements([(o, 0, ())])
        _B, H = _field(L, P, np.array([c["M"]]), np.array([c["p"]]))
        Hm = np.array(c["H"])
        s = np.abs(Hm).max()
        worst_k = max(worst_k, np.abs(H[0] - Hm).max() / s)
        _b, Ho = OR.Poly(c["verts"], c["faces"]).BH(c["p"], c["M"])
        if c.get("near_minus_z"):
            worst_oz = max(worst_oz, np.abs(Ho - Hm).max() / s)
        else:
            worst_o = max(worst_o, np.abs(Ho - Hm).max() / s)
    assert worst_k <= 1e-11
    # documents the reference forms' loss on these elements
    assert worst_o > 1e4 * worst_k
    assert worst_oz > 1e-5


SLIVER_THICKNESS = [1.0, 1e-2, 1e-4, 1e-6, 1e-8]
SLIVER_OBS = np.array([[20.0, 5.0, 3.0], [-15.0, 8.0, -4.0], [5.0, 5.0, 12.0], [0.3, 0.2, 0.05]])
SLIVER_FACES = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
BLOCK_FACES = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]


def sliver_verts(t):
    return [[0, 0, 0], [10, 0, 0], [0, 10, 0], [3, 3, t]]


def sliver_model(t):
    """A sliver tetrahedron (apex t mm above its base) and a solid block, as
    the reference's flatness sweep builds them (tests/test_sliver_tet_asm.py:39-53)."""
    sliver = rad.ObjPolyhdr(sliver_verts(t), [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]], [0, 0, 1])
    rad.MatApl(sliver, rad.MatLin([0.06, 0.17], 1.2))
    block = rad.ObjRecMag([6, 6, 12], [8, 8, 6], [0, 0, 1])
    rad.MatApl(block, rad.MatLin([0.06, 0.17], 1.2))
    return rad.ObjCnt([sliver, block])


def sliver_mp50(t, M_sliver, M_block, pts):
    """H of the sliver and the block at pts (50-digit face-charge closed form),
    and the sliver's self block (columns = H at its centroid for M = e_k)."""
    sys_path = os.path.join(HERE, "golden")
    import sys
    if sys_path not in sys.path:
        sys.path.insert(0, sys_path)
    from make_mp_golden import H_mp
    bv = [[6 + sx * 4, 6 + sy * 4, 12 + sz * 3] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    H = np.array([[float(a) + float(b) for a, b in zip(H_mp(sliver_verts(t), SLIVER_FACES, M_sliver, p),
                                                       H_mp(bv, BLOCK_FACES, M_block, p))] for p in pts])
    c = np.mean(np.array(sliver_verts(t), dtype=float), axis=0)
    Qs = np.array([[float(x) for x in H_mp(sliver_verts(t), SLIVER_FACES, e, c)] for e in np.eye(3)]).T
    return H, Qs


@pytest.mark.parametrize("t", SLIVER_THICKNESS)
def test_sliver_tetrahedron_stays_finite_and_exact(L, t):
    """A sliver down to 1e-8 mm (the reference's flatness sweep): interaction
    blocks and fields finite; the sliver's self block and H at the sweep's
    points within 1e-10 of 50-digit values in FP64 (5e-5, the reference's
    tolerance, in FP32); the oracle agrees while it can (t >= 1e-4).  Below
    that the oracle -- the reference's face-frame form, randomization off --
    loses the self block (Q_zz -> 0 instead of -1 at t = 1e-6)."""
    top = sliver_model(t)
    P, _ = models.packed(top)
    Q = _assemble(L, P)
    assert np.isfinite(Q).all()
    H50, Q50 = sliver_mp50(t, P["M"][0], P["M"][1], SLIVER_OBS)
    assert np.abs(Q[0:3, 0:3] - Q50).max() <= 1e-10
    for fp32, tol in ((0, 1e-10), (1, 5e-5)):
        B, H = _field(L, P, P["M"], SLIVER_OBS, fp32, kinds=bool(fp32))
        assert np.isfinite(B).all() and np.isfinite(H).all()
        assert np.abs(H - H50).max() <= tol * np.abs(H50).max()
    if t >= 1e-4:
        Qo, _jit = models.oracle_assemble(P)
        assert np.abs(Q - Qo).max() <= 5e-5 * np.abs(Qo).max()


def test_mixed_assembly_far_blocks_in_fp32(L):
    """The mixed assembly's integration choice (host build of block_Q_mixed):
    blocks whose collocation point lies within two circumradii of the source
    element are integrated in FP64 (bit-identical to the FP64 assembly), the
    rest in FP32.  On C-yoke rows a far block errs by up to ~1e-3 of its own
    size, but every error stays below 1e-7 of the row's scale, the size of
    FP32 storage rounding."""
    from radiacuda_amd import workloads as W
    rad.UtiDelAll()
    _top, iron, _coil, _info = W.c_yoke()
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    n = len(items)
    desc, keep = _abi.model_desc(P)
    cen = np.array([np.mean(it[0].verts, axis=0) for it in items])
    rad_ = np.array([np.sqrt(((it[0].verts - c) ** 2).sum(axis=1).max()) for it, c in zip(items, cen)])
    for i in (0, 4321, 7999):
        o64 = np.zeros((3, 3 * n))
        omx = np.zeros((3, 3 * n))
        assert L.rhc_assemble(ctypes.byref(desc), ctypes.c_int64(i), ctypes.c_int64(i + 1), ctypes.c_int32(0),
                              o64.ctypes.data_as(dp)) == 0
        assert L.rhc_assemble(ctypes.byref(desc), ctypes.c_int64(i), ctypes.c_int64(i + 1), ctypes.c_int32(2),
                              omx.ctypes.data_as(dp)) == 0
        d = np.abs(omx - o64).reshape(3, n, 3).max(axis=(0, 2))
        s = np.abs(o64).reshape(3, n, 3).max(axis=(0, 2))
        near = np.linalg.norm(cen - cen[i], axis=1) < 2.0 * rad_
        assert near[i] and (d[near] == 0).all()
        assert (d[~near] / s[~near]).max() > 1e-5
        assert d.max() <= 1e-7 * np.abs(o64).max()
    rad.UtiDelAll()
