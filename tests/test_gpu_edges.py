"""Edge cases of the device paths through the C ABI: empty models and point
sets, point counts that do not fill a workgroup, a single element, and
argument errors (rejected on the host before any kernel runs)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from tests import models
from tests.test_gpu_configs import _host_field

pytestmark = pytest.mark.gpu
RH_ERR_ARG = 1  # include/radhip.h


@pytest.fixture(autouse=True)
def _rand_off():
    OR.set_randomization(False)


def test_empty_model_and_empty_point_set():
    print("case: empty model, empty point set", flush=True)
    empty = O.pack_elements([])
    with N.Model(empty) as m:
        B, H = m.field(np.zeros((0, 3)), np.random.default_rng(0).uniform(-1, 1, (10, 3)))
        assert not B.any() and not H.any()
        Q = m.assemble(N.RH_FP64)
        res, M, _H = Q.relax(np.zeros((0, 16)), np.zeros((0, 3)), np.zeros((0, 3)), 1e-6, 10)
        assert res.converged and M.shape == (0, 3)
        Q.close()
    ids = models.tet_mesh((2, 1, 1), 3.0, seed=1)
    P, _ = models.packed(rad.ObjCnt(ids))
    with N.Model(P) as m:
        B, H = m.field(P["M"], np.zeros((0, 3)))
        assert B.shape == (0, 3) and H.shape == (0, 3)


@pytest.mark.parametrize("npts", [1, 63, 64, 65, 255, 257, 4097])
def test_point_counts_that_do_not_fill_a_workgroup(npts):
    """Mesh form (conforming tets) and per-element path (blocks + general
    polyhedron), FP64, against the host build of the same integrals."""
    print("case: %d points" % npts, flush=True)
    rng = np.random.default_rng(npts)
    pts = rng.uniform(-5, 15, (npts, 3))
    mesh = rad.ObjCnt(models.tet_mesh((3, 2, 2), 3.0, seed=3))
    blocks = models.mixed_model(seed=2, sym=False)
    for top in (mesh, blocks):
        P, _ = models.packed(top)
        with N.Model(P) as m:
            B, H = m.field(P["M"], pts)
        Bh, Hh = _host_field(P, P["M"], pts)
        scale = np.abs(Bh).max()
        assert np.abs(B - Bh).max() <= 1e-12 * scale
        assert np.abs(H - Hh).max() <= 1e-12 * np.abs(Hh).max()


def test_single_element_relaxation():
    print("case: single element", flush=True)
    t = rad.ObjPolyhdr([[0, 0, 0], [4, 0, 0], [0, 5, 0], [1, 1, 3]], [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]],
                       [0, 0, 1])
    rad.MatApl(t, rad.MatLin([0.3, 0.1], 1.0))
    k = rad.RlxPre(t)
    it = rad._reg[k]
    Q64 = it.model.assemble(N.RH_FP64)
    Qs = Q64.export()
    K = it.mats[0, 4:13].reshape(3, 3)
    Mref = np.linalg.solve(np.eye(3) - K @ Qs, K @ it.hext[0] + it.mats[0, 13:16])
    for method in (9, 11):
        res, M, _H = Q64.relax(it.mats, it.hext, np.zeros((1, 3)), 1e-13, 100, method=method)
        assert res.converged
        assert np.abs(M[0] - Mref).max() <= 1e-12 * np.abs(Mref).max()
    Q64.close()
    rad.UtiDel(k)


def test_argument_errors_are_rejected():
    print("case: argument errors", flush=True)
    L = N.lib()
    ids = models.tet_mesh((1, 1, 1), 3.0, seed=1)
    P, _ = models.packed(rad.ObjCnt(ids))
    dp = ctypes.POINTER(ctypes.c_double)
    M = np.ascontiguousarray(P["M"])
    pts = np.zeros((4, 3))
    B = np.zeros((4, 3))
    with N.Model(P) as m:
        # negative point count, missing outputs, missing inputs
        assert L.rh_field(m.h, M.ctypes.data_as(dp), pts.ctypes.data_as(dp), -1, B.ctypes.data_as(dp), None,
                          N.RH_FP64) == RH_ERR_ARG
        assert L.rh_field(m.h, M.ctypes.data_as(dp), pts.ctypes.data_as(dp), 4, None, None,
                          N.RH_FP64) == RH_ERR_ARG
        assert L.rh_field(m.h, M.ctypes.data_as(dp), None, 4, B.ctypes.data_as(dp), None,
                          N.RH_FP64) == RH_ERR_ARG
        assert L.rh_field(m.h, None, pts.ctypes.data_as(dp), 4, B.ctypes.data_as(dp), None,
                          N.RH_FP64) == RH_ERR_ARG
        assert b"argument" in L.rh_last_error()
        # the model is still usable afterwards
        Bm, _H = m.field(M, pts + 7.0)
        assert np.isfinite(Bm).all()
    with pytest.raises(N.RadHipError):
        N.Model(P, device=N.lib().rh_device_count() + 3)
