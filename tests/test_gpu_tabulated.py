"""Relaxation with a tabulated isotropic law (MatSatIsoTab) on the GPU: the
hybrid undulator of the reference's third example (steel poles with the
example's M(H) table, NdFeB blocks, three symmetry planes), against the FP64
Newton fixed point of the oracle's matrix and the oracle's method-4 relaxation
with the same node table."""
import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu


def test_tabulated_undulator_relaxation():
    print("case: hybrid undulator with tabulated steel", flush=True)
    OR.set_randomization(False)
    top, _steel = models.hybrid_undulator()
    P, items = models.packed(top)
    n = len(items)
    k = rad.RlxPre(top)
    it = rad._reg[k]
    assert (it.mats[:, 0] == 3).any() and len(it.tab) > 0
    res = rad.RlxAuto(k, 1e-11, 3000)
    print("gpu relax (mixed + refinement)", res, flush=True)
    M = np.stack([o.M for (o, _k, _s) in items])
    Qo, _mask = models.oracle_assemble(P)
    OR.set_tables(it.tab)
    try:
        Mo, _H, iters, stats = OR.relax_m4(Qo.astype(np.float32), it.mats, it.hext, 1e-7, 20000)
    finally:
        OR.set_tables(None)
    print("oracle m4", iters, stats, flush=True)
    Mn = models.fp64_fixed_point(Qo, it.mats, it.hext, M0=Mo, tab=it.tab)
    scale = np.abs(Mn).max()
    d_ref = np.abs(M - Mn).max() / scale
    d_ora = np.abs(M - Mo).max() / scale
    print("refined vs FP64 fixed point %.3e, vs oracle m4 %.3e" % (d_ref, d_ora), flush=True)
    assert res[0] <= 1e-11
    assert d_ref <= 1e-7
    # FP32 dot products of the oracle's Gauss-Seidel move its fixed point
    assert d_ora <= 1e-4
    # FP64 storage, no refinement needed
    Q64 = it.model.assemble(N.RH_FP64)
    r64, M64, _H64 = Q64.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-12, 3000, tab=it.tab)
    assert r64.converged
    assert np.abs(M64 - Mn).max() <= 1e-8 * scale
    Q64.close()
    # peak field on axis (the example prints Bz at the origin)
    Bz = rad.Fld(top, "bz", [0, 0, 0])
    # the origin lies on the symmetry planes, i.e. on face planes of the
    # blocks: the oracle takes it with Radia's default length randomization
    Bo, _Ho, _bad = models.oracle_field(P, Mn, np.zeros((1, 3)))
    print("Bz(0) device %.9f T, oracle on the Newton state %.9f T" % (Bz, Bo[0, 2]), flush=True)
    assert abs(Bz - Bo[0, 2]) <= 1e-7 * abs(Bo[0, 2])
    rad.UtiDel(k)
    # the example's own call: Solve(und, 0.0003, 1000)
    r = rad.Solve(top, 0.0003, 1000)
    assert r[0] <= 3e-4
    assert abs(rad.Fld(top, "bz", [0, 0, 0]) - Bo[0, 2]) <= 1e-3 * abs(Bo[0, 2])


def test_tabulated_record_outside_table_is_refused():
    print("case: tabulated record without table", flush=True)
    top, _steel = models.hybrid_undulator(n_periods=1)
    k = rad.RlxPre(top)
    it = rad._reg[k]
    n = len(it.items)
    with pytest.raises(N.RadHipError, match="outside the table"):
        it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-6, 10, tab=None)
    bad = it.mats.copy()
    bad[bad[:, 0] == 3, 1] = len(it.tab) // 3  # first node past the end
    with pytest.raises(N.RadHipError, match="outside the table"):
        it.matrix.relax(bad, it.hext, np.zeros((n, 3)), 1e-6, 10, tab=it.tab)
    rad.UtiDel(k)
