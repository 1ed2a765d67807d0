"""Symmetric-block storage of the interaction matrix.

Without transform copies and row transforms every off-diagonal block is the
field tensor of source j at collocation point i, a symmetric 3x3 matrix,
traceless when the point lies outside the source.  Such matrices are stored
as 5 planes per element row (xx, xy, xz, yy, yz; zz = -(xx + yy)), or 6 when a
collocation point lies inside another element; rh_matrix_opts.dense_rows
keeps the row-major 3x3 blocks the reference stores (radintrc.h:126).

Checks against the row-major storage of the same model:
* exported rows: every stored entry equal to the row-major one (to rounding),
  and zz = -(xx + yy) off the row-major zz by no more than that block's
  computed trace (zero in exact arithmetic; below 1e-13 of the block in FP64);
* the relaxed state of a nonlinear steel mesh (prec 1e-11) within 1e-7 of the
  FP64 fixed point (Newton) of the row-major FP64 matrix, as the row-major
  solve is, and the linear solve within 1e-11 of a direct solve of the
  exported system;
* the transposed (adjoint) solve against a direct solve of the transposed
  system;
* the layout chosen: 5 planes for a tet mesh, 6 when a block's collocation
  point lies inside another element, row-major under a symmetry transform;
* streamed rows: bit-identical to the resident matrix.
"""
import numpy as np
import pytest

from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu


def _block_scale_rel(A, B, n):
    d = np.abs(A - B).reshape(n, 3, n, 3).max(axis=(1, 3))
    s = np.abs(B).reshape(n, 3, n, 3).max(axis=(1, 3))
    return (d / np.maximum(s, 1e-300)).max()


def _steel_mesh(seed=7):
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=seed, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    return iron, src


def test_layout_choice_and_export():
    print("case: symmetric layout choice and export, tet mesh", flush=True)
    rad.UtiDelAll()
    iron, _src = _steel_mesh()
    P, items = models.packed(iron)
    n = len(items)
    with N.Model(P) as model:
        for prec, tol in ((N.RH_FP64, 4.5e-16), (N.RH_MIXED, 1.2e-7)):
            Qs = model.assemble(prec)
            Qd = model.assemble(prec, dense_rows=True)
            fs, fd = Qs.format(), Qd.format()
            print("prec %d: symmetric layout %s, dense %s" % (prec, fs, fd), flush=True)
            assert fs[0] == 5 and fs[1] % 256 == 0 and fs[1] >= n
            assert fd[0] == 0
            As, Ad = Qs.export(), Qd.export()
            Bs, Bd = As.reshape(n, 3, n, 3), Ad.reshape(n, 3, n, 3)
            scale = np.abs(Bd).max(axis=(1, 3))[:, None, :, None]
            # the dense storage holds both off-diagonal entries: they agree
            asym = (np.abs(Bd - Bd.transpose(0, 3, 2, 1)) / scale).max()
            # every stored entry is the dense one; zz = -(xx + yy) differs from
            # the dense zz by the dense block's computed trace (zero in exact
            # arithmetic outside the source), which bounds the blocks' rounding
            keep = np.ones((3, 3), bool)
            keep[2, 2] = False
            d = np.abs(Bs - Bd) / scale
            tr = np.abs(np.einsum("iaja->ij", Bd))[:, None, :, None] / scale
            tr_max = tr[:, 0, :, 0][~np.eye(n, dtype=bool)].max()
            e_entries = d.transpose(0, 2, 1, 3)[:, :, keep].max()
            e_zz = (d[:, 2, :, 2] - tr[:, 0, :, 0]).max()
            print("  dense block asymmetry %.3e, stored entries %.3e, zz beyond the dense trace %.3e, "
                  "dense trace %.3e of the block scale" % (asym, e_entries, e_zz, tr_max), flush=True)
            assert asym <= 4 * tol and e_entries <= tol and e_zz <= 2 * tol
            if prec == N.RH_FP64:
                assert tr_max <= 1e-13
            # self blocks are held in FP64 in both
            for e in range(n):
                assert np.array_equal(As[3 * e:3 * e + 3, 3 * e:3 * e + 3], Ad[3 * e:3 * e + 3, 3 * e:3 * e + 3])
            Qs.close()
            Qd.close()


def test_relaxation_matches_dense_storage():
    print("case: nonlinear and linear relaxation, symmetric vs dense storage", flush=True)
    rad.UtiDelAll()
    iron, src = _steel_mesh(seed=11)
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    with N.Model(P) as model:
        hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, src))), model.colloc())
        # FP64 fixed point of the row-major FP64 matrix (Newton polish)
        Q64 = model.assemble(N.RH_FP64, dense_rows=True)
        A64 = Q64.export()
        Q64.close()
        Mfp = None
        for prec in (N.RH_FP64, N.RH_MIXED):
            Qs = model.assemble(prec)
            Qd = model.assemble(prec, dense_rows=True)
            rs, Ms, Hs = Qs.relax(mats, hext, np.zeros((n, 3)), 1e-11, 3000)
            rd, Md, Hd = Qd.relax(mats, hext, np.zeros((n, 3)), 1e-11, 3000)
            if Mfp is None:
                Mfp = models.fp64_fixed_point(A64, mats, hext, M0=Md)
            es = np.abs(Ms - Mfp).max() / np.abs(Mfp).max()
            ed = np.abs(Md - Mfp).max() / np.abs(Mfp).max()
            print("prec %d: iterations %d / %d, vs the FP64 fixed point: symmetric %.3e, dense %.3e"
                  % (prec, rs.iters, rd.iters, es, ed), flush=True)
            # residual 1e-11 times the conditioning of chi ~ 2000 steel
            assert rs.converged and rd.converged
            assert es <= 1e-7 and ed <= 1e-7
            Qs.close()
            Qd.close()
        # linear material: the relaxed state solves the exported system
        lin = np.zeros_like(mats)
        lin[:, 0] = 0
        lin[:, 1] = 4.0
        lin[:, 13:16] = np.random.default_rng(3).uniform(-1, 1, (n, 3))
        Q = model.assemble(N.RH_FP64)
        assert Q.format()[0] == 5
        A = Q.export()
        r, M, _H = Q.relax(lin, hext, np.zeros((n, 3)), 1e-13, 3000)
        K = 4.0 * np.eye(3 * n)
        Md = np.linalg.solve(np.eye(3 * n) - K @ A, K @ hext.ravel() + lin[:, 13:16].ravel())
        e = np.abs(M.ravel() - Md).max() / np.abs(Md).max()
        print("linear: %d iterations, vs direct solve %.3e" % (r.iters, e), flush=True)
        assert r.converged and e <= 1e-11
        # transposed system (adjoint of the linear solve)
        rhs = np.random.default_rng(4).normal(size=(n, 3))
        res, lam = Q.relax_adjoint(lin, rhs, 1e-13, 500)
        lam_d = np.linalg.solve(np.eye(3 * n) - A.T @ K.T, rhs.ravel())
        e = np.abs(lam.ravel() - lam_d).max() / np.abs(lam_d).max()
        print("adjoint: %d iterations, vs direct solve %.3e" % (res.iters, e), flush=True)
        assert res.converged and e <= 1e-11
        Q.close()


def test_six_planes_when_a_point_is_inside_another_element():
    print("case: overlapping elements -> 6 planes", flush=True)
    rad.UtiDelAll()
    rng = np.random.default_rng(8)
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=9)
    # a block overlapping the mesh: its centre and several tet centroids lie
    # inside elements other than their own (at generic positions)
    ids.append(rad.ObjRecMag([5.3, 6.1, 3.7], [3.1, 2.9, 3.3], list(rng.uniform(-1, 1, 3))))
    top = rad.ObjCnt(ids)
    P, items = models.packed(top)
    n = len(items)
    with N.Model(P) as model:
        for prec, tol in ((N.RH_FP64, 4.5e-16), (N.RH_MIXED, 1.2e-7)):
            Qs = model.assemble(prec)
            Qd = model.assemble(prec, dense_rows=True)
            print("prec %d: layout %s" % (prec, Qs.format()), flush=True)
            assert Qs.format()[0] == 6
            As, Ad = Qs.export(), Qd.export()
            tr = np.einsum("iaja->ij", As.reshape(n, 3, n, 3))
            np.fill_diagonal(tr, 0.0)
            assert (np.abs(tr) > 0.5).any()  # the traced blocks are there
            # six planes store every entry; the off-diagonal pairs are averaged
            Bd = Ad.reshape(n, 3, n, 3)
            asym = (np.abs(Bd - Bd.transpose(0, 3, 2, 1)).max(axis=(1, 3)) / np.abs(Bd).max(axis=(1, 3))).max()
            err = _block_scale_rel(As, Ad, n)
            print("  dense asymmetry %.3e, six planes vs dense %.3e" % (asym, err), flush=True)
            assert err <= 0.5 * asym + tol
            # streamed rows of a 6-plane matrix (the trace is met in a tile)
            Qt = model.assemble(prec, resident_fraction=0.0, allow_streaming=True)
            assert Qt.format()[0] == 6 and Qt.layout()[0] == 0
            assert np.array_equal(Qt.export(), As)
            for Q in (Qs, Qd, Qt):
                Q.close()


def test_row_major_under_a_symmetry_transform():
    print("case: symmetry transform keeps the row-major layout", flush=True)
    rad.UtiDelAll()
    top = models.mixed_model(seed=1, sym=True)
    P, _items = models.packed(top)
    with N.Model(P) as model:
        Q = model.assemble(N.RH_FP64)
        assert Q.format()[0] == 0
        Q.close()


def test_streamed_symmetric_rows_are_bit_identical():
    print("case: streamed symmetric-block rows", flush=True)
    rad.UtiDelAll()
    iron, src = _steel_mesh(seed=13)
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    with N.Model(P) as model:
        hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, src))), model.colloc())
        for prec in (N.RH_FP64, N.RH_MIXED):
            ref = None
            for frac in (1.0, 0.5, 0.0):
                Q = model.assemble(prec, resident_fraction=frac, allow_streaming=True)
                assert Q.format()[0] == 5
                r, M, H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-9, 3000)
                rows = Q.export(0, 3 * n)
                print("prec %d fraction %.1f: layout %s, %d iterations" % (prec, frac, Q.layout(), r.iters),
                      flush=True)
                if ref is None:
                    ref = (M, H, rows, r.iters)
                else:
                    assert np.array_equal(M, ref[0]) and np.array_equal(H, ref[1])
                    assert np.array_equal(rows, ref[2]) and r.iters == ref[3]
                Q.close()
