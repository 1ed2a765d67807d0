"""The C-shaped dipole yoke at full size on the GPU (8,000 jittered
tetrahedra, MatSatIsoFrm steel, two racetrack coils, 24,000 x 24,000 matrix,
Solve prec 1e-4, B on the 100^3 gap grid): the workload the benchmark times.

Parity against the oracle is split by the one mechanism that makes the
reference's CPU path inaccurate on this mesh:

* four elements have a face whose outward normal lies within 1e-7 of -z
  without being -z (|1 + n_z| = 2e-8 and 2e-10).  The reference's face
  rotation (radvlpgn.cpp:357-371) then takes the axis-aligned frame and drops
  the tilt; the oracle, restating it, is off by up to 5e-3 of those elements'
  fields.  They are checked against 50-digit values (tests/golden/mp50_yoke.json);
* every other element is checked against the oracle: <= 1e-9 of the field
  (measured 4.4e-10: the oracle's edge and face-frame forms lose up to 4e-8
  of a distant element's field);
* the whole model's field (a fixed synthetic magnetization) is checked
  against its sum over all elements at 50 digits: <= 1e-12, where the
  oracle is off by 1e-7 (tests/golden/mp50_yoke_full.json).

The tight FP64 solve is checked by its physics residual M - M_mat(H_ext + Q M),
recomputed from the field path and the oracle's material law; the
mixed-precision path (FP32 storage, FP32 far-field assembly and field
integrals, FP64 refinement) against the FP64 oracle field, <= 1e-6, both at
the configuration's prec 1e-4 (against the FP64 solve stopped there) and at
the fixed point.
"""
import json
import os

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


def oracle_currents(cur, pts):
    """B of the packed current sources by the oracle (lab frame)."""
    B = np.zeros((len(pts), 3))
    for k in range(len(cur["kind"])):
        p = cur["par"][k]
        Mt, V, s = cur["cM"][k], cur["cV"][k], cur["cS"][k]
        for i, x in enumerate(pts):
            xl = np.linalg.solve(Mt, x - V)
            if cur["kind"][k] == 0:
                b, _ = OR.reccur_BH(xl, p[0:3], p[3:6], p[6:9])
            else:
                b, _ = OR.arccur_BH(xl, p[0:3], p[3], p[4], p[5], p[6], p[7], p[8], int(p[9]))
            B[i] += s * (Mt @ b)
    return B


def physics_residual(mats, M, H):
    """RMS over elements of |M - M_mat(H)| (oracle material law)."""
    r = np.array([M[e] - OR.mat_M(mats[e], H[e]) for e in range(len(M))])
    return float(np.sqrt((r * r).sum(axis=1).mean()))


def test_c_yoke_full_size_parity():
    OR.set_randomization(False)
    print("case: C-yoke 8000 tetrahedra, FP64 and mixed, 1M points", flush=True)
    top, iron, coil, info = W.c_yoke()
    P, items = models.packed(iron)
    n = len(items)
    assert n == 8000
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    cur = O.pack_currents(list(O.walk(rad._reg, coil)))
    model = N.Model(P)
    X = model.colloc()
    hext = N.current_field(cur, X)
    pts = W.gap_grid(info, 100)
    rng = np.random.default_rng(2)
    samp = np.sort(rng.choice(len(pts), 64, replace=False))

    # FP64 solve at the configuration's prec, then to the fixed point
    Q = model.assemble(N.RH_FP64)
    res, M, H = Q.relax(mats, hext, np.zeros((n, 3)), 1e-4, 1000)
    assert res.converged and res.misfit <= 1e-4
    res_t, Mt, Ht = Q.relax(mats, hext, M, 1e-11, 5000, keep_m=True)
    Q.close()
    # the physics residual of the tight state, H recomputed with the field path
    Hc = model.field(Mt, X)[1] + hext
    r_t = physics_residual(mats, Mt, Hc)
    r_4 = physics_residual(mats, M, model.field(M, X)[1] + hext)
    dM = np.abs(M - Mt).max() / np.abs(Mt).max()
    print("C-yoke FP64: %d it to 1e-4 (misfit %.2e), fixed point after %d more; physics residual "
          "%.2e at the fixed point, %.2e at prec 1e-4; max |M - M*| / max |M*| = %.2e"
          % (res.iters, res.misfit, res_t.iters, r_t, r_4, dM), flush=True)
    # prec 1e-4 (the configuration's) bounds the implicit misfit, not the
    # physics residual: at chi ~ 2000 the state is still ~10 % (max |dM|)
    # from the fixed point there, and B ~ 5e-4 of its scale
    checks = [("tight solve misfit", res_t.misfit, 1e-11),
              ("physics residual at the fixed point", r_t, 1e-8)]

    # field on the full grid (device path), FP64 and mixed
    B64 = model.field(M, pts)[0]
    Bt = model.field(Mt, pts[samp])[0]
    checks.append(("B at prec 1e-4 vs the fixed point", np.abs(B64[samp] - Bt).max() / np.abs(Bt).max(), 2e-3))
    ps = pts[samp]

    # (b) every element but the tilted-face ones, against the oracle
    flagged = models.tilted_minus_z(items)
    gold = json.load(open(os.path.join(HERE, "golden", "mp50_yoke.json")))["c_yoke"]
    assert sorted(flagged.tolist()) == sorted(gold["flagged"])
    Mz = Mt.copy()
    Mz[flagged] = 0.0
    Bz = model.field(Mz, ps)[0]
    Boz, _Hoz, bad = models.oracle_field(P, Mz, ps)
    assert not bad.any()
    e_unfl = np.abs(Bz - Boz).max() / np.abs(Boz).max()

    # (c) the tilted-face elements (and the eight elements where the oracle
    # loses most otherwise) against 50-digit values
    gp = np.array(gold["points"])
    Mf = np.zeros_like(Mt)
    Mf[flagged] = Mt[flagged]
    Bf = model.field(Mf, gp)[0]
    Bf_mp = np.zeros_like(Bf)
    for e in gold["flagged"]:
        Bf_mp += np.einsum("pab,b->pa", np.array(gold["Q"][str(e)]), Mt[e])
    Bfull_gp = model.field(Mt, gp)[0]
    e_fl = np.abs(Bf - Bf_mp).max() / np.abs(Bfull_gp).max()
    e_blk = 0.0
    for e in gold["flagged"] + gold["unflagged_worst"]:
        Qmp = np.array(gold["Q"][str(e)])
        for c in range(3):
            Me = np.zeros_like(Mt)
            Me[e, c] = 1.0
            Hg = model.field(Me, gp)[1]
            d = np.abs(Hg - Qmp[:, :, c]).max(axis=1) / np.abs(Qmp).max(axis=(1, 2))
            e_blk = max(e_blk, d.max())
    # the whole model with a fixed synthetic magnetization against its field
    # summed over all 8,000 elements at 50 digits (tests/golden/mp50_yoke_full.json)
    full = json.load(open(os.path.join(HERE, "golden", "mp50_yoke_full.json")))["c_yoke"]
    Ms = synthetic_M(n)
    fp = np.array(full["points"])
    Bmp = np.array(full["B"])
    e_mp = np.abs(model.field(Ms, fp)[0] - Bmp).max() / np.abs(Bmp).max()
    e_mp_oracle = np.abs(models.oracle_field(P, Ms, fp)[0] - Bmp).max() / np.abs(Bmp).max()
    print("C-yoke whole-model field vs 50 digits: FP64 device path %.2e, oracle %.2e"
          % (e_mp, e_mp_oracle), flush=True)
    checks.append(("whole-model field vs 50 digits", e_mp, 1e-12))
    Bo_full = models.oracle_field(P, Mt, ps)[0]
    e_all = np.abs(Bt - Bo_full).max() / np.abs(Bo_full).max()
    print("C-yoke FP64 field vs oracle: all elements %.2e; without the %d tilted-face elements "
          "%.2e; tilted-face elements vs 50 digits %.2e of the field, per block %.2e"
          % (e_all, len(flagged), e_unfl, e_fl, e_blk), flush=True)
    # the other elements: the oracle's edge and face-frame forms lose up to
    # 4e-8 of a distant element's field, 4.4e-10 of this field in all; the
    # whole-model 50-digit sum above pins the device path at 3e-15
    checks += [("unflagged elements vs oracle", e_unfl, 1e-9), ("tilted-face elements vs 50 digits", e_fl, 1e-12),
               ("blocks vs 50 digits", e_blk, 1e-10)]

    # (e) mixed path (FP32 storage, FP32 far-field assembly and field
    # integrals, FP64 refinement) against the FP64 oracle: at the
    # configuration's prec 1e-4 against the oracle field of the FP64 solve
    # stopped at the same prec, and solved to the fixed point against the
    # oracle field of the FP64 fixed point
    Qm = model.assemble(N.RH_MIXED)
    res_m, Mm, _Hm = Qm.relax(mats, hext, np.zeros((n, 3)), 1e-4, 1000)
    assert res_m.converged and res_m.misfit <= 1e-4
    res_mt, Mmt, _Hmt = Qm.relax(mats, hext, Mm, 1e-11, 5000, keep_m=True, refine=3)
    Qm.close()
    Bm = model.field(Mm, pts, precision=N.RH_MIXED)[0]
    Bmt = model.field(Mmt, ps, precision=N.RH_MIXED)[0]
    coil_o = oracle_currents(cur, ps)
    coil_d = N.current_field(cur, ps)
    Bo_prec = models.oracle_field(P, M, ps)[0] + coil_o
    Bo_tot = Bo_full + coil_o
    scale = np.abs(Bo_tot).max()
    e_mixed = np.abs(Bm[samp] + coil_d - Bo_prec).max() / scale
    e_mixed_t = np.abs(Bmt + coil_d - Bo_tot).max() / scale
    e_fp64 = np.abs(B64[samp] + coil_d - Bo_prec).max() / scale
    print("C-yoke B (elements + coils) vs the FP64 oracle: mixed at prec 1e-4 %.2e (%d it, refined %d), "
          "FP64 at prec 1e-4 %.2e; mixed at the fixed point %.2e (misfit %.1e after %d more it, "
          "refined %d); mixed vs FP64 on the grid %.2e"
          % (e_mixed, res_m.iters, res_m.refined, e_fp64, e_mixed_t, res_mt.misfit, res_mt.iters,
             res_mt.refined, np.abs(Bm - B64).max() / np.abs(B64).max()), flush=True)
    checks += [("mixed B at prec 1e-4 vs FP64 oracle", e_mixed, 1e-6),
               ("mixed B at the fixed point vs FP64 oracle", e_mixed_t, 1e-6)]
    model.close()
    for name, val, bound in checks:
        assert val <= bound, "%s: %.3e > %.1e" % (name, val, bound)
