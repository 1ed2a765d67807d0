"""The benchmark configurations other than the timed one, at full size, on the
GPU: parity against the oracle where it finishes in seconds (sampled rows and
points) and size-independent properties elsewhere.

* iron cube: 512 blocks (ObjDivMag 8x8x8), MatSatIsoFrm steel, four-bar
  ObjRecCur loop, B on 10,000 points of a tilted line through the centre.
* large yoke: 40,000 tetrahedra, 120,000 x 120,000 matrix, FP32 storage
  (57.6 GB) with FP64 refinement and FP64 storage (115 GB), prec 1e-5.
* Halbach ring: 20,000 hexahedral polyhedra, MatLin(0.06/0.17, 1.2 T), B at
  2,000,000 points in the bore, FP32 vs FP64 field kernels.
"""
import json
import os
import time

import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from radiacuda_amd import workloads as W
from tests import models
from tests.golden.make_mp_golden import synthetic_M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(autouse=True)
def _rand_off():
    OR.set_randomization(False)


def _ext_field(coil_ids, X):
    items = list(O.walk(rad._reg, coil_ids))
    return N.current_field(O.pack_currents(items), X)


def _host_rows(P, i):
    import ctypes
    import os
    from radiacuda_amd import _abi
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                   "radiacuda_amd", "_rhcheck.so"))
    desc, keep = _abi.model_desc(P)
    n = len(P["type"])
    out = np.zeros((3, 3 * n))
    assert lib.rhc_assemble(ctypes.byref(desc), ctypes.c_int64(i), ctypes.c_int64(i + 1), ctypes.c_int32(0),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return out


def _host_field(P, M, pts):
    import ctypes
    import os
    from radiacuda_amd import _abi
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                   "radiacuda_amd", "_rhcheck.so"))
    desc, keep = _abi.model_desc(P)
    dp = ctypes.POINTER(ctypes.c_double)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    B = np.zeros_like(pts)
    H = np.zeros_like(pts)
    assert lib.rhc_field(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                         ctypes.c_int64(len(pts)), ctypes.c_int32(0), B.ctypes.data_as(dp),
                         H.ctypes.data_as(dp)) == 0
    return B, H


def test_iron_cube_full_size():
    print("case: iron cube 512 blocks", flush=True)
    top, iron, coil, pts = W.iron_cube()
    P, items = models.packed(iron)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P)
    X = model.colloc()
    hext = _ext_field(coil, X)
    # FP64 matrix vs the oracle (every block)
    Q = model.assemble(N.RH_FP64)
    Qs = Q.export()
    Qo, _jit = models.oracle_assemble(P)
    assert models.block_rel(Qs, Qo, n) <= 1e-10
    t = time.perf_counter()
    res, M, _H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-11, 3000)
    t_solve = time.perf_counter() - t
    assert res.converged
    # FP64 fixed point of the oracle's system (Newton polish from the GPU
    # result): the RMS residual 1e-11 T times the iron's conditioning
    Mn = models.fp64_fixed_point(Qo, mats, hext, M0=M)
    print("iron cube: %d iterations to misfit %.2e, vs FP64 fixed point %.2e"
          % (res.iters, res.misfit, np.abs(M - Mn).max() / np.abs(Mn).max()), flush=True)
    assert np.abs(M - Mn).max() <= 1e-7 * np.abs(Mn).max()
    # field of the relaxed state at the 10,000 line points: elements + coil
    t = time.perf_counter()
    B, _Hf = model.field(M, pts)
    t_fld = time.perf_counter() - t
    Bo, _Ho, bad = models.oracle_field(P, M, pts)
    assert not bad.any()
    assert np.abs(B - Bo).max() <= 1e-10 * np.abs(Bo).max()
    # drop-in API end to end (mixed storage + FP64 refinement) at the
    # configuration's prec 1e-4, against the FP64-storage solve stopped at the
    # same prec: the same iteration on a matrix that differs by FP32 rounding
    r = rad.Solve(top, 1e-4, 1000)
    Bapi = np.array([rad.Fld(top, "b", list(p)) for p in pts[::500]])
    r4, M4, _H4 = Q.relax(mats, hext, np.zeros((n, 3)), 1e-4, 1000)
    B4 = model.field(M4, pts[::500])[0] + _ext_field(coil, pts[::500])
    Bconv = B[::500] + _ext_field(coil, pts[::500])
    scale = np.abs(B4).max()
    print("iron cube: solve %.3f s (%d it), field %.3f s; Solve(1e-4) %s; API vs FP64 at 1e-4 %.2e; "
          "at 1e-4 vs converged (1e-10) %.2e"
          % (t_solve, res.iters, t_fld, r, np.abs(Bapi - B4).max() / scale,
             np.abs(B4 - Bconv).max() / scale), flush=True)
    assert r[0] <= 1e-4
    assert np.abs(Bapi - B4).max() <= 1e-3 * scale
    Q.close()
    model.close()


def test_large_yoke_120k_matrix():
    print("case: large yoke 40000 tets", flush=True)
    top, iron, coil, info = W.c_yoke(ny=50)
    P, items = models.packed(iron)
    n = len(items)
    assert n == 40000
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P)
    X = model.colloc()
    hext = N.current_field(O.pack_currents(list(O.walk(rad._reg, coil))), X)
    rows = [0, 12345, 39999]
    out = {}
    for prec in (N.RH_MIXED, N.RH_FP64):
        t = time.perf_counter()
        Q = model.assemble(prec)
        t_asm = time.perf_counter() - t
        R = np.concatenate([Q.export(3 * i, 3) for i in rows])
        t = time.perf_counter()
        res, M, _H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-5, 1000)
        t_solve = time.perf_counter() - t
        print("prec %d: assembly %.3f s, solve %.3f s, %d iterations, misfit %.2e, refined %d"
              % (prec, t_asm, t_solve, res.iters, res.misfit, res.refined), flush=True)
        assert res.converged and res.misfit <= 1e-5
        # on to a tight prec from there (the mixed matrix with its FP64
        # refinement): both reach the FP64 system's fixed point
        t = time.perf_counter()
        rt, Mt, _H = Q.relax(mats, hext, M, 1e-9, 4000, keep_m=True, refine=3)
        print("prec %d: on to misfit %.2e in %d more iterations (%.2f s), refined %d"
              % (prec, rt.misfit, rt.iters, time.perf_counter() - t, rt.refined), flush=True)
        assert rt.converged
        out[prec] = (R, M, Mt)
        Q.close()
    # sampled rows: against the host build of the same integrals, against the
    # oracle on every column but the tilted-face elements (a face within 1e-7
    # of -z: the reference's face rotation drops its tilt, radvlpgn.cpp:357-371),
    # and those columns against 50-digit values (tests/golden/mp50_yoke.json)
    R64, M64, M64t = out[N.RH_FP64]
    R32, M32, M32t = out[N.RH_MIXED]
    Rh = np.concatenate([_host_rows(P, i) for i in rows])
    om = OR.prepare_model(P)
    Ro = np.concatenate([OR.assemble(om, rows=(i, i + 1)) for i in rows])
    flagged = models.tilted_minus_z(items)
    gold = json.load(open(os.path.join(HERE, "golden", "mp50_yoke.json")))["large_yoke"]
    assert sorted(flagged.tolist()) == sorted(gold["flagged"]) and gold["rows"] == rows
    fc = np.zeros(3 * n, bool)
    for e in flagged:
        fc[3 * e:3 * e + 3] = True
    scale = np.abs(Ro).max()
    e_host = np.abs(R64 - Rh).max() / scale
    e_unfl = np.abs(R64 - Ro)[:, ~fc].max() / scale
    e_fl = e_blk = 0.0
    for r, i in enumerate(rows):
        assert np.abs(X[i] - np.array(gold["colloc"][r])).max() <= 1e-13 * np.abs(X[i]).max()
        rs = np.abs(Ro[3 * r:3 * r + 3]).max()
        for e in flagged:
            Qmp = np.array(gold["Q"][str(e)][r])
            Qg = R64[3 * r:3 * r + 3, 3 * e:3 * e + 3]
            e_fl = max(e_fl, np.abs(Qg - Qmp).max() / rs)
            e_blk = max(e_blk, np.abs(Qg - Qmp).max() / np.abs(Qmp).max())
    print("large yoke rows: vs host build %.2e, vs oracle without the %d tilted-face columns %.2e "
          "(all columns %.2e); tilted-face blocks vs 50 digits %.2e of the row, %.2e per block"
          % (e_host, len(flagged), e_unfl, np.abs(R64 - Ro).max() / scale, e_fl, e_blk), flush=True)
    assert e_host <= 1e-12
    assert e_unfl <= 1e-10
    assert e_fl <= 1e-12
    assert e_blk <= 1e-10
    assert np.abs(R32 - R64).max() <= 1e-7 * scale
    # both solves stop at prec 1e-5 on slightly different paths; the states
    # agree to the stopping tolerance's reach (max |dM| ~ 4e-4 of max |M| at
    # chi ~ 2000), not to the precision of either matrix
    dM = np.abs(M32 - M64).max() / np.abs(M64).max()
    # at prec 1e-9 the refined mixed state is the FP64 one to the precision
    # target: magnetization and the field it makes in the gap
    dMt = np.abs(M32t - M64t).max() / np.abs(M64t).max()
    gpts = W.gap_grid(info, 20)[::97]
    B64t = model.field(M64t, gpts)[0]
    dBt = np.abs(model.field(M32t, gpts)[0] - B64t).max() / np.abs(B64t).max()
    print("large yoke: refined mixed vs FP64 solve, max |dM|/max|M| = %.2e at prec 1e-5; at prec 1e-9 "
          "%.2e, gap field %.2e" % (dM, dMt, dBt), flush=True)
    assert dM <= 1e-3
    assert dMt <= 1e-5 and dBt <= 1e-6
    # field at gap points: host build; oracle without the tilted-face
    # elements; those elements against 50-digit values
    pts = W.gap_grid(info, 20)[::97]
    B, _H = model.field(M64, pts)
    Bh, _Hh = _host_field(P, M64, pts)
    Mz = M64.copy()
    Mz[flagged] = 0.0
    Bz = model.field(Mz, pts)[0]
    Boz, _Ho, _bad = models.oracle_field(P, Mz, pts)
    gp = np.array(gold["points"])
    Mf = np.zeros_like(M64)
    Mf[flagged] = M64[flagged]
    Bf = model.field(Mf, gp)[0]
    Bf_mp = np.zeros_like(Bf)
    for e in flagged:
        Bf_mp += np.einsum("pab,b->pa", np.array(gold["Q"][str(e)][len(rows):]), M64[e])
    e_fl = np.abs(Bf - Bf_mp).max() / np.abs(model.field(M64, gp)[0]).max()
    scale = np.abs(Boz).max()
    print("large yoke field: vs host build %.2e, vs oracle without the tilted-face elements %.2e; "
          "tilted-face elements vs 50 digits %.2e of the field"
          % (np.abs(B - Bh).max() / np.abs(Bh).max(), np.abs(Bz - Boz).max() / scale, e_fl), flush=True)
    # the whole model with a fixed synthetic magnetization against its field
    # summed over all 40,000 elements at 50 digits
    full = json.load(open(os.path.join(HERE, "golden", "mp50_yoke_full.json")))["large_yoke"]
    Ms = synthetic_M(n)
    fp = np.array(full["points"])
    Bmp = np.array(full["B"])
    e_mp = np.abs(model.field(Ms, fp)[0] - Bmp).max() / np.abs(Bmp).max()
    Bmp_o = models.oracle_field(P, Ms, fp)[0]
    print("large yoke whole-model field vs 50 digits: FP64 device path %.2e, oracle %.2e"
          % (e_mp, np.abs(Bmp_o - Bmp).max() / np.abs(Bmp).max()), flush=True)
    assert e_mp <= 1e-12
    assert np.abs(B - Bh).max() <= 1e-12 * np.abs(Bh).max()
    # the oracle's own error on the other elements is 5e-10 of this field (its
    # edge and face-frame forms lose up to 4e-8 of a distant element's field;
    # the 50-digit sum above pins the device path at 1e-14)
    assert np.abs(Bz - Boz).max() <= 2e-9 * scale
    assert e_fl <= 1e-12
    model.close()


def test_halbach_field_2m_points():
    print("case: halbach 20000 hexahedra, 2M points", flush=True)
    top, pts = W.halbach()
    k = rad.RlxPre(top)
    t = time.perf_counter()
    r = rad.RlxAuto(k, 1e-8, 1000)
    t_solve = time.perf_counter() - t
    assert r[0] <= 1e-8
    it = rad._reg[k]
    items = it.items
    P = it.packed
    M = np.stack([o.M for (o, _k, _s) in items])
    model = it.model
    times = {}
    Bs = {}
    for prec in (N.RH_FP64, N.RH_MIXED):
        t = time.perf_counter()
        B, _H = model.field(M, pts, precision=prec)
        times[prec] = time.perf_counter() - t
        Bs[prec] = B
    scale = np.abs(Bs[N.RH_FP64]).max()
    err32 = np.abs(Bs[N.RH_MIXED] - Bs[N.RH_FP64]).max() / scale
    print("halbach: solve %.3f s; field FP64 %.2f s, FP32 %.2f s (2M points); FP32 vs FP64 %.2e"
          % (t_solve, times[N.RH_FP64], times[N.RH_MIXED], err32), flush=True)
    assert err32 <= 1e-6
    rng = np.random.default_rng(3)
    samp = rng.choice(len(pts), 48, replace=False)
    Bo, _Ho, _bad = models.oracle_field(P, M, pts[samp])
    assert np.abs(Bs[N.RH_FP64][samp] - Bo).max() <= 1e-10 * np.abs(Bo).max()
    rad.UtiDel(k)
