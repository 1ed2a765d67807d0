"""Method 11 (Newton-Krylov) on the GPU: the same fixed points as the
Anderson solver and the FP64 references, on linear, tanh-law, tabulated and
high-susceptibility (iron cube, 512 blocks) models.  The misfit of method 11
is the physics residual RMS |M_mat(H) - M| and its iteration count the number
of matrix passes (radgpurlx.h:78-86)."""
import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from radiacuda_amd import workloads as W
from tests import models

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _oracle_off():
    OR.set_randomization(False)


def _physics_residual(Q, mats, hext, M, tab=None):
    n = len(mats)
    H = (Q @ M.ravel()).reshape(n, 3) + hext
    F = np.stack([models.mat_M_and_D(mats[e], H[e], tab)[0] for e in range(n)]) - M
    return np.sqrt((F ** 2).sum() / n)


def test_nk_linear_matches_direct_solve():
    print("case: newton-krylov linear anisotropic", flush=True)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=6, M=[0, 0, 1.0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatLin([0.06, 0.17], 1.2))
    src = rad.ObjRecCur([6, -9, 4], [6, 4, 10], [0.0, 0.0, 3.0])
    top = rad.ObjCnt([iron, src])
    P, items = models.packed(iron)
    n = len(items)
    k = rad.RlxPre(top)
    it = rad._reg[k]
    Q64 = it.model.assemble(N.RH_FP64)
    Qs = Q64.export()
    A = np.eye(3 * n)
    rhs = np.zeros(3 * n)
    for e in range(n):
        K = it.mats[e, 4:13].reshape(3, 3)
        A[3 * e:3 * e + 3, :] -= K @ Qs[3 * e:3 * e + 3, :]
        rhs[3 * e:3 * e + 3] = K @ it.hext[e] + it.mats[e, 13:16]
    Mref = np.linalg.solve(A, rhs).reshape(n, 3)
    res, M, _H = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-13, 500, method=11)
    print("nk linear: %d matvecs, %d newton, %d krylov, misfit %.2e"
          % (res.matvecs, res.newton, res.krylov, res.misfit), flush=True)
    assert res.converged and res.misfit <= 1e-13
    assert res.iters == res.matvecs
    assert np.abs(M - Mref).max() <= 1e-11 * np.abs(Mref).max()
    # a linear system: each Newton step cuts |F| by its forcing term (0.5 at
    # the first step, then quadratically smaller)
    assert res.newton <= 6
    Q64.close()
    # the drop-in API with meth=11 (mixed storage + FP64 refinement)
    r = rad.RlxAuto(k, 1e-12, 1000, 11)
    Mapi = np.stack([o.M for (o, _k, _s) in items])
    assert r[0] <= 1e-12
    assert np.abs(Mapi - Mref).max() <= 1e-10 * np.abs(Mref).max()
    rad.UtiDel(k)


def test_nk_nonlinear_matches_fixed_point():
    """chi ~ 2000 iron driven by a coil: method 11 reaches the FP64 Newton
    fixed point of the stored system, as method 9 does."""
    print("case: newton-krylov nonlinear iron", flush=True)
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=7, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    top = rad.ObjCnt([iron, src])
    P, items = models.packed(iron)
    n = len(items)
    k = rad.RlxPre(top)
    it = rad._reg[k]
    Q64 = it.model.assemble(N.RH_FP64)
    Qs = Q64.export()
    r9, M9, _ = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 5000)
    r11, M11, H11 = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 5000, method=11)
    print("method 9: %d matvecs, misfit %.2e; method 11: %d matvecs (%d newton, %d krylov), misfit %.2e"
          % (r9.matvecs, r9.misfit, r11.matvecs, r11.newton, r11.krylov, r11.misfit), flush=True)
    assert r11.converged and r11.misfit <= 1e-12
    Mn = models.fp64_fixed_point(Qs, it.mats, it.hext, M0=M11)
    scale = np.abs(Mn).max()
    assert np.abs(M11 - Mn).max() <= 1e-8 * scale
    assert np.abs(M11 - M9).max() <= 1e-7 * scale
    # the reported misfit is the physics residual of the returned state
    assert _physics_residual(Qs, it.mats, it.hext, M11) <= 2e-12
    # H output: the total field at the collocation points
    Hs = (Qs @ M11.ravel()).reshape(n, 3) + it.hext
    assert np.abs(H11 - Hs).max() <= 1e-10 * np.abs(Hs).max()
    Q64.close()
    # mixed storage + FP64 refinement through the API
    r = rad.RlxAuto(k, 1e-11, 5000, 11)
    Mapi = np.stack([o.M for (o, _k, _s) in items])
    Qo, _jit = models.oracle_assemble(P)
    Mo = models.fp64_fixed_point(Qo, it.mats, it.hext, M0=Mapi)
    print("api meth 11:", r, "vs FP64 fixed point %.2e" % (np.abs(Mapi - Mo).max() / scale), flush=True)
    assert r[0] <= 1e-11
    assert np.abs(Mapi - Mo).max() <= 1e-7 * scale
    rad.UtiDel(k)


def test_nk_tabulated_undulator():
    print("case: newton-krylov tabulated steel undulator", flush=True)
    top, _steel = models.hybrid_undulator()
    P, items = models.packed(top)
    n = len(items)
    k = rad.RlxPre(top)
    it = rad._reg[k]
    Q64 = it.model.assemble(N.RH_FP64)
    Qs = Q64.export()
    r9, M9, _ = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 5000, tab=it.tab)
    r11, M11, _ = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 5000, tab=it.tab, method=11)
    print("method 9: %d matvecs; method 11: %d matvecs (%d newton)" % (r9.matvecs, r11.matvecs, r11.newton),
          flush=True)
    assert r11.converged
    Mn = models.fp64_fixed_point(Qs, it.mats, it.hext, M0=M11, tab=it.tab)
    scale = np.abs(Mn).max()
    assert np.abs(M11 - Mn).max() <= 1e-9 * scale
    assert np.abs(M9 - Mn).max() <= 1e-8 * scale
    Q64.close()
    rad.UtiDel(k)


def test_nk_iron_cube_full_size():
    print("case: newton-krylov iron cube 512 blocks", flush=True)
    top, iron, coil, pts = W.iron_cube()
    P, items = models.packed(iron)
    n = len(items)
    mats, tab = O.pack_materials(items)
    model = N.Model(P)
    X = model.colloc()
    hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, coil))), X)
    Q = model.assemble(N.RH_FP64)
    Qs = Q.export()
    r9, M9, _ = Q.relax(mats, hext, np.zeros((n, 3)), 1e-10, 5000)
    r11, M11, _ = Q.relax(mats, hext, np.zeros((n, 3)), 1e-10, 5000, method=11)
    print("iron cube: method 9 %d matvecs; method 11 %d matvecs (%d newton, %d krylov)"
          % (r9.matvecs, r11.matvecs, r11.newton, r11.krylov), flush=True)
    assert r11.converged
    Mn = models.fp64_fixed_point(Qs, mats, hext, M0=M11)
    scale = np.abs(Mn).max()
    assert np.abs(M11 - Mn).max() <= 1e-7 * scale
    Q.close()
    model.close()


def test_nk_option_checks():
    print("case: newton-krylov option checks", flush=True)
    ids = models.tet_mesh((2, 2, 2), 4.0, jitter=0.1, seed=4, M=[0, 0, 1.0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatLin([0.06, 0.17], 1.2))
    k = rad.RlxPre(iron)
    it = rad._reg[k]
    n = len(it.items)
    with pytest.raises(N.RadHipError, match="krylov_m"):
        it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-6, 10, method=11, krylov_m=64)
    with pytest.raises(N.RadHipError, match="method"):
        it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-6, 10, method=7)
    # restart lengths 2 and 60 reach the same state (restarted GMRES)
    r2, M2, _H = it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 2000, method=11, krylov_m=2)
    r60, M60, _H = it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 2000, method=11)
    print("krylov_m 2: %d matvecs; 60: %d matvecs" % (r2.matvecs, r60.matvecs), flush=True)
    assert r2.converged and r60.converged
    assert np.abs(M2 - M60).max() <= 1e-10 * np.abs(M60).max()
    rad.UtiDel(k)
