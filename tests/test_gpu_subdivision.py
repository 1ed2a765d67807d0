"""A polyhedron subdivided by ObjDivMag and a thick polygon, relaxed and
evaluated on the GPU through the drop-in API: the relaxed magnetization
against the direct FP64 solve of the linear system with the device matrix,
that matrix against the oracle's, and the field against the oracle."""
import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu


def test_divided_tetrahedron_and_thick_polygon_relax():
    print("case: ObjDivMag tetrahedron + ObjThckPgn prism, linear material", flush=True)
    rad.UtiDelAll()
    OR.set_randomization(False)
    t = rad.ObjPolyhdr([[0, 0, 0], [30, 0, 0], [0, 30, 0], [0, 0, 30]],
                       [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]], [0, 0, 1])
    rad.ObjDivMag(t, [[4, 1.5], 3, 3])
    p = rad.ObjThckPgn(-8.0, 6.0, [[0, 0], [20, 0], [25, 12], [5, 18]], "x", [0, 1, 0])
    iron = rad.ObjCnt([t, p])
    rad.MatApl(iron, rad.MatLin([0.2, 0.05], 0.8))
    src = rad.ObjRecCur([10, 10, 45], [40, 40, 8], [0, 0, 2.0])
    top = rad.ObjCnt([iron, src])
    P, items = models.packed(iron)
    n = len(items)
    assert n == 17  # 16 pieces of the tetrahedron and the prism
    k = rad.RlxPre(top)
    it = rad._reg[k]
    Q64 = it.model.assemble(N.RH_FP64)
    Qs = Q64.export()
    Qo, _mask = models.oracle_assemble(P)
    assert np.abs(Qs - Qo).max() <= 1e-10 * np.abs(Qo).max()
    A = np.eye(3 * n)
    rhs = np.zeros(3 * n)
    for e in range(n):
        K = it.mats[e, 4:13].reshape(3, 3)
        A[3 * e:3 * e + 3, :] -= K @ Qs[3 * e:3 * e + 3, :]
        rhs[3 * e:3 * e + 3] = K @ it.hext[e] + it.mats[e, 13:16]
    Mref = np.linalg.solve(A, rhs).reshape(n, 3)
    Q64.close()
    r = rad.RlxAuto(k, 1e-12, 2000)
    M = np.stack([o.M for (o, _k, _s) in items])
    print("relax:", r, "vs direct %.2e" % (np.abs(M - Mref).max() / np.abs(Mref).max()), flush=True)
    assert r[0] <= 1e-12
    assert np.abs(M - Mref).max() <= 1e-9 * np.abs(Mref).max()
    pts = np.random.default_rng(5).uniform(-20, 45, (500, 3))
    B = np.array(rad.Fld(iron, "b", pts.tolist()))
    Bo, _Ho, _bad = models.oracle_field(P, M, pts)
    assert np.abs(B - Bo).max() <= 1e-10 * np.abs(Bo).max()
    rad.UtiDel(k)
