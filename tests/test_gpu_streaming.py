"""Out-of-core interaction matrix: rows beyond the device budget live in pinned
host memory and stream through two device tiles in every matrix pass.

* The relaxed magnetization is bit-identical whatever fraction of the rows is
  resident (all streamed, half, all resident), as each row is reduced the same
  way wherever it lives; exported rows are identical too.
* UtiGpuFallback policy: 'break' refuses a matrix beyond the budget with the
  reference's Error602 message text; 'gpu_streaming' streams it and reports
  the backend (tests/test_gpu_fallback_policy.py of the reference).
"""
import numpy as np
import pytest

from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from radiacuda_amd import workloads as W
from tests import models

pytestmark = pytest.mark.gpu


def test_streamed_solve_is_bit_identical():
    print("case: streamed matrix, C-yoke, resident fractions 0 / 0.5 / 1", flush=True)
    top, iron, coil, info = W.c_yoke()
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P)
    hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, coil))), model.colloc())
    ref = None
    for frac in (None, 0.0, 0.5, 1.0):
        kw = {} if frac is None else dict(resident_fraction=frac, allow_streaming=True)
        Q = model.assemble(N.RH_MIXED, **kw)
        res_rows, str_rows = Q.layout()
        rows = Q.export(3 * (n - 2), 6)
        r, M, H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-4, 1000)
        print("fraction %s: %d resident rows, %d streamed, %d iterations, misfit %.3e"
              % (frac, res_rows, str_rows, r.iters, r.misfit), flush=True)
        assert res_rows + str_rows == 3 * n
        if frac is not None:
            # the resident part is a whole number of 64-element groups
            res = n if frac == 1.0 else int(frac * n) // 64 * 64
            assert str_rows == 3 * (n - res)
        if ref is None:
            ref = (M, H, rows, r.iters)
        else:
            assert np.array_equal(M, ref[0]) and np.array_equal(H, ref[1])
            assert np.array_equal(rows, ref[2]) and r.iters == ref[3]
        Q.close()
    model.close()


def test_fallback_policy(monkeypatch):
    print("case: fallback policy break / gpu_streaming", flush=True)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=51, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    top = rad.ObjCnt([iron, src])
    try:
        rad.UtiGpuFallback("break")
        k = rad.RlxPre(top)
        r0 = rad.RlxAuto(k, 1e-9, 1000)
        M0 = np.stack([o.M for (o, _k, _s) in rad._reg[k].items])
        assert rad.UtiAsmLastBackend() == "gpu"
        rad.UtiDel(k)
        monkeypatch.setenv("RADHIP_MATRIX_DEVICE_BYTES", "100000")
        with pytest.raises(RuntimeError, match="fallback policy is 'break'"):
            rad.RlxPre(top)
        rad.UtiGpuFallback("gpu_streaming")
        k = rad.RlxPre(top)
        # the reference reports 'gpu' for a streamed assembly too; the C ABI
        # tells the two apart (2 = device with streamed rows)
        assert rad.UtiAsmLastBackend() == "gpu" and N.last_backend(0) == 2
        assert rad._reg[k].matrix.layout()[1] > 0
        r1 = rad.RlxAuto(k, 1e-9, 1000)
        M1 = np.stack([o.M for (o, _k, _s) in rad._reg[k].items])
        assert r1[3] == r0[3] and np.array_equal(M1, M0)
        rad.UtiDel(k)
        # 'cpu' streams too (there is no host assembly here)
        rad.UtiGpuFallback("cpu")
        k = rad.RlxPre(top)
        assert N.last_backend(0) == 2 and rad._reg[k].matrix.layout()[1] > 0
        rad.UtiDel(k)
    finally:
        rad.UtiGpuFallback("cpu")


def test_streamed_nonlinear_solve_reaches_fixed_point():
    """Half the rows streamed from pinned host memory: the nonlinear (tanh
    steel) solve reaches the FP64 fixed point of the matrix it stores (Newton
    reference on the exported matrix) and that of the oracle's matrix, as the
    in-core solve does (radgpurlx.cu:258-459 is the reference's streamed
    pass)."""
    from oracle import oracle as OR
    print("case: streamed nonlinear solve vs FP64 fixed point", flush=True)
    OR.set_randomization(False)
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=7, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    with N.Model(P) as model:
        hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, src))), model.colloc())
        Qo, _jit = models.oracle_assemble(P)
        for prec in (N.RH_FP64, N.RH_MIXED):
            Q = model.assemble(prec, resident_fraction=0.5, allow_streaming=True)
            res_rows, str_rows = Q.layout()
            assert str_rows == 3 * (n - n // 2 // 64 * 64)
            r, M, _H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-11, 3000)
            assert r.converged and r.misfit <= 1e-11
            Qs = Q.export()
            Ms = models.fp64_fixed_point(Qs if prec == N.RH_FP64 else Qo, mats, hext, M0=M)
            err = np.abs(M - Ms).max() / np.abs(Ms).max()
            print("streamed %s: %d resident / %d streamed rows, %d iterations, vs FP64 fixed point %.2e"
                  % ("fp64" if prec == N.RH_FP64 else "mixed+refine", res_rows, str_rows, r.iters, err),
                  flush=True)
            assert err <= 1e-7
            Q.close()
