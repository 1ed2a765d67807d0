"""GPU adjoint kernels through the C ABI.

* dL/dM: B is linear in M, so dL/dM_j = sum_p (dB_p/dM_j)^T g_p with the
  columns dB_p/dM_j taken from the oracle field of element j at unit M
  (<= 1e-10 relative), with symmetry copies, blocks and polyhedra.
* dL/d(vertex), dL/d(point): against the host build of the same dual-number
  math (<= 1e-10) and against central finite differences of the oracle
  (<= 1e-6; the host tests pin the math to 1e-7).
* Relaxation adjoint: lam solves (I - Q^T K^T) lam = rhs on the exported
  stored matrix (residual <= 1e-10).
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _abi
from radiacuda_amd import _native as N
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIBCHK = os.path.join(os.path.dirname(HERE), "radiacuda_amd", "_rhcheck.so")
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(autouse=True)
def _rand_off():
    OR.set_randomization(False)


def _rot(rng):
    q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_adjoint_dM_matches_oracle_columns():
    print("case: adjoint dM mixed model with symmetries", flush=True)
    rng = np.random.default_rng(21)
    top = models.mixed_model(seed=2, sym=True)
    P, items = models.packed(top)
    n = len(items)
    pts = rng.uniform(-12, 16, (150, 3))
    g = rng.normal(size=pts.shape)
    model = N.Model(P)
    gM, _gV, _gX = model.field_adjoint(None, pts, g, field="b", want_m=True)
    ref = np.zeros((n, 3))
    for k in range(3):
        M = np.zeros((n, 3))
        for j in range(n):
            M[:] = 0.0
            M[j, k] = 1.0
            B, _H, _bad = models.oracle_field(P, M, pts)
            ref[j, k] = np.sum(g * B)
    assert np.isfinite(gM).all()
    assert np.abs(gM - ref).max() <= 1e-10 * np.abs(ref).max()
    # H weights: the inside term drops out
    gMh, _a, _b = model.field_adjoint(None, pts, g, field="h", want_m=True)
    refh = np.zeros((n, 3))
    for j in range(n):
        for k in range(3):
            M = np.zeros((n, 3))
            M[j, k] = 1.0
            _B, H, _bad = models.oracle_field(P, M, pts)
            refh[j, k] = np.sum(g * H)
    assert np.abs(gMh - refh).max() <= 1e-10 * np.abs(refh).max()
    model.close()


def _chk_grad(P, M, pts, g):
    L = ctypes.CDLL(LIBCHK)
    desc, keep = _abi.model_desc(P)
    pts, M, g = (np.ascontiguousarray(a, dtype=np.float64) for a in (pts, M, g))
    gV = np.zeros((len(P["verts"]), 3))
    gX = np.zeros_like(pts)
    assert L.rhc_adj_grad(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                          ctypes.c_int64(len(pts)), g.ctypes.data_as(dp), gV.ctypes.data_as(dp),
                          gX.ctypes.data_as(dp)) == 0
    return gV, gX


def test_adjoint_vertex_and_point_gradients():
    print("case: adjoint dV dX tets + hexahedra", flush=True)
    rng = np.random.default_rng(23)
    ids = models.tet_mesh((2, 2, 1), 4.0, jitter=0.1, seed=24)
    for k in range(3):
        ids.append(models.hex_block([12.0 + 5 * k, 2.0, 1.0], [3.0, 2.0, 4.0], rng.uniform(-1, 1, 3),
                                    _rot(rng)))
    P, items = models.packed(rad.ObjCnt(ids))
    P["verts"] = np.ascontiguousarray(P["verts"] @ _rot(rng).T)
    n = len(items)
    M = rng.uniform(-1, 1, (n, 3))
    pts = rng.uniform(-10, 25, (300, 3))
    g = rng.normal(size=pts.shape)
    model = N.Model(P)
    assert model.vertex_count() == len(P["verts"])
    _gM, gV, gX = model.field_adjoint(M, pts, g, field="h", want_m=False, want_v=True, want_x=True)
    cV, cX = _chk_grad(P, M, pts, g)
    assert np.abs(gV - cV).max() <= 1e-10 * np.abs(cV).max()
    assert np.abs(gX - cX).max() <= 1e-10 * np.abs(cX).max()
    # finite differences of the oracle field in a few vertex coordinates (tets)
    h = 1e-5
    for v, c in [(0, 0), (5, 1), (9, 2), (17, 0)]:
        Vp, Vm = P["verts"].copy(), P["verts"].copy()
        Vp[v, c] += h
        Vm[v, c] -= h
        Pp, Pm = dict(P), dict(P)
        Pp["verts"], Pm["verts"] = Vp, Vm
        fd = (np.sum(g * OR.field(OR.prepare_model(Pp), M, pts)[1]) -
              np.sum(g * OR.field(OR.prepare_model(Pm), M, pts)[1])) / (2 * h)
        assert abs(fd - gV[v, c]) <= 1e-6 * np.abs(gV).max()
    model.close()


def test_adjoint_symmetric_model_point_gradient_vs_fd():
    """Symmetry copies: dL/dx of the GPU adjoint against finite differences
    of the GPU forward field; dL/dvertex along a rigid translation of each
    element against finite differences of the oracle."""
    print("case: adjoint symmetric tets", flush=True)
    rng = np.random.default_rng(25)
    ids = models.tet_mesh((2, 1, 1), 4.0, origin=(1, 2, 3), jitter=0.1, seed=26)
    a = rad.ObjCnt(ids)
    rad.TrfZerPara(a, [0, 0, 0], [0.3, 0.2, 1])
    rad.TrfOrnt(a, rad.TrfRot([1, 1, 1], [0.2, 0.3, 1.0], 0.4))
    P, items = models.packed(rad.ObjCnt([a]))
    n = len(items)
    M = rng.uniform(-1, 1, (n, 3))
    pts = rng.uniform(-12, 14, (80, 3))
    g = rng.normal(size=pts.shape)
    model = N.Model(P)
    _gM, gV, gX = model.field_adjoint(M, pts, g, field="b", want_m=False, want_v=True, want_x=True)
    h = 1e-5
    for p in range(0, len(pts), 7):
        for c in range(3):
            xp, xm = pts.copy(), pts.copy()
            xp[p, c] += h
            xm[p, c] -= h
            fd = (np.sum(g[p] * model.field(M, xp[p:p + 1])[0]) -
                  np.sum(g[p] * model.field(M, xm[p:p + 1])[0])) / (2 * h)
            assert abs(fd - gX[p, c]) <= 1e-6 * np.abs(gX).max()
    vo = P["vert_off"]
    for e in range(0, n, 3):
        d = rng.normal(size=3)
        dirs = np.zeros_like(P["verts"])
        dirs[vo[e]:vo[e + 1]] = d
        Pp, Pm = dict(P), dict(P)
        Pp["verts"], Pm["verts"] = P["verts"] + h * dirs, P["verts"] - h * dirs
        fd = (np.sum(g * OR.field(OR.prepare_model(Pp), M, pts)[0]) -
              np.sum(g * OR.field(OR.prepare_model(Pm), M, pts)[0])) / (2 * h)
        an = np.sum(gV * dirs)
        assert abs(fd - an) <= 1e-6 * max(abs(an), np.abs(gV).max())
    model.close()


def test_relax_adjoint_solves_transposed_system():
    print("case: relax adjoint linear anisotropic", flush=True)
    rng = np.random.default_rng(27)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=28, M=[0, 0, 1.0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatLin([0.06, 0.17], 1.2))
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P)
    for prec in (N.RH_FP64, N.RH_MIXED):
        Q = model.assemble(prec)
        Qs = Q.export()
        rhs = rng.normal(size=(n, 3))
        res, lam = Q.relax_adjoint(mats, rhs, 1e-13, 500)
        K = np.zeros((3 * n, 3 * n))
        for e in range(n):
            K[3 * e:3 * e + 3, 3 * e:3 * e + 3] = mats[e, 4:13].reshape(3, 3)
        A = np.eye(3 * n) - Qs.T @ K.T
        r = A @ lam.ravel() - rhs.ravel()
        print("adjoint relax prec %d: iters %d misfit %.3e residual %.3e"
              % (prec, res.iters, res.misfit, np.abs(r).max()), flush=True)
        assert res.converged
        assert np.abs(r).max() <= 1e-10 * np.abs(rhs).max()
        Q.close()
    model.close()


def _undulator_problem(n_blocks, n_points, seed=0):
    import torch
    from radiacuda_amd import adjoint as A
    from radiacuda_amd import objects as O
    from radiacuda_amd import workloads as W
    top, info = W.undulator(n_blocks, n_points=n_points, seed=seed)
    items = list(O.walk(rad._reg, top))
    P = O.pack_elements(items)
    recs = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    dev = torch.device("cuda")
    geom = A.Geometry(P)
    verts = torch.tensor(P["verts"], dtype=torch.float64, device=dev)
    Mr = torch.tensor(recs[:, 13:16], dtype=torch.float64, device=dev)
    K = torch.tensor(recs[:, 4:13], dtype=torch.float64, device=dev)
    pts = torch.tensor(info["points"], dtype=torch.float64, device=dev)
    target = torch.tensor(info["target"], dtype=torch.float64, device=dev)

    def loss(Mr_, verts_):
        M = A.relax_linear(Mr_, verts_, geom, K, prec=1e-14)
        B = A.field(M, verts_, pts, geom)
        return ((B[:, 1] - target) ** 2).sum()

    return P, recs, info, Mr, verts, loss


def _fd_grads(loss, Mr, verts, gMr, gV, coords_m, coords_v, hm=1e-6, hv=1e-6):
    import torch
    worst = 0.0
    scale_m = gMr.abs().max().item()
    scale_v = gV.abs().max().item()
    with torch.no_grad():
        for (e, c) in coords_m:
            Mp, Mm = Mr.clone(), Mr.clone()
            Mp[e, c] += hm
            Mm[e, c] -= hm
            fd = (loss(Mp, verts) - loss(Mm, verts)).item() / (2 * hm)
            err = abs(fd - gMr[e, c].item()) / scale_m
            worst = max(worst, err)
        for (v, c) in coords_v:
            Vp, Vm = verts.clone(), verts.clone()
            Vp[v, c] += hv
            Vm[v, c] -= hv
            fd = (loss(Mr, Vp) - loss(Mr, Vm)).item() / (2 * hv)
            err = abs(fd - gV[v, c].item()) / scale_v
            worst = max(worst, err)
    return worst


def test_undulator_gradient_small_vs_fp64_reference_and_fd():
    """40 blocks: dL/dMr against the closed-form FP64 gradient from the oracle's
    dense matrix and field columns; dL/dvertex against central differences."""
    print("case: undulator 40 blocks autograd", flush=True)
    import torch
    P, recs, info, Mr, verts, loss = _undulator_problem(40, 64)
    Mr.requires_grad_(True)
    verts.requires_grad_(True)
    L = loss(Mr, verts)
    L.backward()
    gMr, gV = Mr.grad.detach(), verts.grad.detach()
    # FP64 reference: M = (I - K Q)^-1 Mr, B = G M
    n = len(recs)
    Qo, _jit = models.oracle_assemble(P)
    K = np.zeros((3 * n, 3 * n))
    for e in range(n):
        K[3 * e:3 * e + 3, 3 * e:3 * e + 3] = recs[e, 4:13].reshape(3, 3)
    A = np.eye(3 * n) - K @ Qo
    pts = info["points"]
    G = np.zeros((3 * len(pts), 3 * n))
    for j in range(n):
        for k in range(3):
            M = np.zeros((n, 3))
            M[j, k] = 1.0
            B, _H, _bad = models.oracle_field(P, M, pts)
            G[:, 3 * j + k] = B.ravel()
    M = np.linalg.solve(A, recs[:, 13:16].ravel())
    By = (G @ M).reshape(-1, 3)[:, 1]
    w = np.zeros((len(pts), 3))
    w[:, 1] = 2 * (By - info["target"])
    ref = np.linalg.solve(A.T, G.T @ w.ravel()).reshape(n, 3)
    print("dL/dMr vs FP64 reference %.3e" % (np.abs(gMr.cpu().numpy() - ref).max() / np.abs(ref).max()))
    assert np.abs(gMr.cpu().numpy() - ref).max() <= 1e-9 * np.abs(ref).max()
    assert abs(L.item() - np.sum((By - info["target"]) ** 2)) <= 1e-9 * L.item()
    rng = np.random.default_rng(5)
    coords_v = [(int(v), int(c)) for v, c in zip(rng.integers(0, len(P["verts"]), 12), rng.integers(0, 3, 12))]
    worst = _fd_grads(loss, Mr.detach(), verts.detach(), gMr, gV, [], coords_v)
    print("dL/dvertex vs central differences %.3e" % worst, flush=True)
    assert worst <= 1e-5


def test_undulator_gradient_full_size_vs_fd():
    """2,000 blocks, 4,096 on-axis points: gradients in 3 magnetization
    components and 24 vertex coordinates per block through the linear solve,
    central differences (h = 1e-6 mm, 1e-6 T) on a seeded sample."""
    print("case: undulator 2000 blocks autograd", flush=True)
    import time
    import torch
    P, recs, info, Mr, verts, loss = _undulator_problem(2000, 4096)
    Mr.requires_grad_(True)
    verts.requires_grad_(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    L = loss(Mr, verts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print("forward %.3f s, backward %.3f s" % (t1 - t, t2 - t1), flush=True)
    gMr, gV = Mr.grad.detach(), verts.grad.detach()
    assert torch.isfinite(gMr).all() and torch.isfinite(gV).all()
    rng = np.random.default_rng(7)
    n = len(recs)
    coords_m = [(int(e), int(c)) for e, c in zip(rng.integers(0, n, 4), rng.integers(0, 3, 4))]
    # blocks near the middle of the device, where the points are
    mid = rng.integers(n // 2 - 60, n // 2 + 60, 4)
    coords_v = [(int(8 * e + rng.integers(0, 8)), int(rng.integers(0, 3))) for e in mid]
    worst = _fd_grads(loss, Mr.detach(), verts.detach(), gMr, gV, coords_m, coords_v)
    print("full size: worst relative error vs central differences %.3e" % worst, flush=True)
    assert worst <= 1e-5
