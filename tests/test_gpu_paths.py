"""GPU parity of the three paths against the CPU oracle (FP64 unless stated).

Tolerances: FP64 field and interaction blocks <= 1e-10 relative to the
block / field scale; FP32 matrix storage <= 2e-7 relative per block;
relaxed magnetization within a small multiple of the requested precision.
"""
import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu


def _pts_off_planes(rng, n, lo, hi):
    # random points; a random point lies on an element plane with probability 0
    return rng.uniform(lo, hi, (n, 3))


def _block_rel(A, B, n):
    d = np.abs(A - B).reshape(n, 3, n, 3).max(axis=(1, 3))
    s = np.abs(B).reshape(n, 3, n, 3).max(axis=(1, 3))
    return (d / np.maximum(s, 1e-300)).max()


@pytest.fixture
def oracle_off():
    OR.set_randomization(False)


def test_field_fp64_matches_oracle(oracle_off):
    print("case: field fp64 mixed model", flush=True)
    top = models.mixed_model(seed=1)
    P, _items = models.packed(top)
    rng = np.random.default_rng(5)
    pts = _pts_off_planes(rng, 700, -12, 16)
    with N.Model(P) as m:
        B, H = m.field(P["M"], pts)
        Bo, Ho, bad = models.oracle_field(P, P["M"], pts)
        assert not bad.any()
        assert np.abs(B - Bo).max() <= 1e-10 * np.abs(Bo).max()
        assert np.abs(H - Ho).max() <= 1e-10 * np.abs(Ho).max()
        # inside elements B = H + M (collocation points of each element); a
        # collocation point on another element's face plane is 0/0 for the
        # unrandomized reference and is compared with the jittered one
        X = m.colloc()
        B2, _ = m.field(P["M"], X)
        Bo2, _, bad = models.oracle_field(P, P["M"], X)
        assert np.isfinite(B2).all()
        s = np.abs(Bo2).max()
        assert np.abs(B2 - Bo2)[~bad].max() <= 1e-10 * s
        if bad.any():
            assert np.abs(B2 - Bo2)[bad].max() <= 1e-6 * s


def test_field_small_point_count_split(oracle_off):
    """Few points and many elements exercise the element-split reduction."""
    print("case: field split", flush=True)
    ids = models.tet_mesh((4, 4, 2), 3.0, jitter=0.1, seed=3)
    top = rad.ObjCnt(ids)
    P, _ = models.packed(top)
    rng = np.random.default_rng(1)
    pts = rng.uniform(-3, 15, (37, 3))
    with N.Model(P) as m:
        B, _H = m.field(P["M"], pts)
    Bo, _, _bad = models.oracle_field(P, P["M"], pts)
    assert np.abs(B - Bo).max() <= 1e-10 * np.abs(Bo).max()


def test_assembly_fp64_matches_oracle(oracle_off):
    print("case: assembly fp64", flush=True)
    top = models.mixed_model(seed=2)
    P, _ = models.packed(top)
    n = len(P["type"])
    with N.Model(P) as m:
        q = m.assemble(N.RH_FP64)
        Q = q.export()
        q.close()
    Qo, jit = models.oracle_assemble(P)
    clean, jittered = models.block_rel(Q, Qo, n, jit)
    assert np.isfinite(Q).all()
    assert clean <= 1e-10
    assert jittered <= 1e-6


def test_assembly_fp32_storage(oracle_off):
    """RH_MIXED: FP32 storage, FP64 self blocks.  With every block integrated
    in FP64 each stored block is the FP64 block rounded (<= 2e-7 of it).  By
    default the far blocks (beyond two circumradii) are integrated in FP32:
    near blocks stay within FP32 rounding, and a far block's FP32 integral
    errs by up to ~1e-3 of its own size, which is ~1e-8 of the row's scale
    (the far field is small); every block is within 3e-7 of its row scale.
    Tetrahedra, blocks and hexahedral general polyhedra (integrated in FP64:
    their FP32 integrals measured no faster on the Halbach ring)."""
    print("case: assembly fp32 storage", flush=True)
    ids = models.tet_mesh((5, 4, 3), 5.0, jitter=0.1, seed=4)
    ids.append(rad.ObjRecMag([-6, 3, 2], [3, 4, 5], [0, 0, 1]))
    ids.append(rad.ObjRecMag([60, 3, 2], [3, 4, 5], [0, 0, 1]))
    for c in ([-8, 14, 6], [-8, 18, 6], [45, -10, 4]):
        ids.append(models.hex_block(c, [3, 4, 5], [0.2, -0.4, 1.0]))
    top = rad.ObjCnt(ids)
    P, items = models.packed(top)
    n = len(P["type"])
    Qo, jit = models.oracle_assemble(P)
    with N.Model(P) as m:
        X = m.colloc()
        q = m.assemble(N.RH_MIXED, fp64_integrals=True)
        Q64i = q.export()
        q.close()
        q = m.assemble(N.RH_MIXED)
        Q = q.export()
        q.close()
    clean, jittered = models.block_rel(Q64i, Qo, n, jit)
    assert clean <= 2e-7 and jittered <= 1e-6
    # near blocks: centre within two circumradii of the source element
    near = np.zeros((n, n), bool)
    for j, (o, _k, _s) in enumerate(items):
        if hasattr(o, "verts"):
            c = np.mean(o.verts, axis=0)
            r = np.sqrt(((o.verts - c) ** 2).sum(axis=1).max())
        else:
            c, r = np.asarray(o.center), 0.5 * np.linalg.norm(o.dims)
        near[:, j] = np.linalg.norm(X - c, axis=1) < 2.0 * r
    d = np.abs(Q - Qo).reshape(n, 3, n, 3).max(axis=(1, 3))
    s = np.abs(Qo).reshape(n, 3, n, 3).max(axis=(1, 3))
    row = np.abs(Qo).reshape(n, 3 * 3 * n).max(axis=1)
    ok = near & ~jit
    assert (d[ok] / s[ok]).max() <= 2e-7
    far_rel = (d / s)[~near & ~jit].max()
    print("mixed assembly: near blocks %d of %d; far blocks %.2e of their size, %.2e of the row"
          % (near.sum(), n * n, far_rel, (d / row[:, None]).max()), flush=True)
    assert (d / row[:, None]).max() <= 3e-7
    assert far_rel > 1e-6  # the far blocks did take the FP32 integrals
    # diagonal blocks are kept in FP64 (FP64 parity with the oracle)
    for i in range(n):
        blk = slice(3 * i, 3 * i + 3)
        assert np.abs(Q[blk, blk] - Qo[blk, blk]).max() <= 1e-10 * np.abs(Qo[blk, blk]).max()


def test_relax_linear_matches_direct_solve(oracle_off):
    """Linear anisotropic material: the relaxed M solves (I - K Q) M = K H + M_r
    on the matrix the device stores (FP32 off-diagonal, FP64 diagonal blocks),
    and the stored matrix agrees with the FP64 oracle to FP32 rounding."""
    print("case: relax linear", flush=True)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=6, M=[0, 0, 1.0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatLin([0.06, 0.17], 1.2))
    src = rad.ObjRecCur([6, -9, 4], [6, 4, 10], [0.0, 0.0, 3.0])
    top = rad.ObjCnt([iron, src])
    P, items = models.packed(iron)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    k = rad.RlxPre(top)
    it = rad._reg[k]
    res = rad.RlxAuto(k, 1e-12, 2000)
    M = np.stack([o.M for (o, _k, _s) in items])
    Qs = it.matrix.export()
    n = len(items)
    Qo, _jit = models.oracle_assemble(P)
    # FP32 storage, far blocks integrated in FP32: within 3e-7 of each row's scale
    row = np.abs(Qo).reshape(n, 9 * n).max(axis=1)
    assert (np.abs(Qs - Qo).reshape(n, 3, n, 3).max(axis=(1, 3)) / row[:, None]).max() <= 3e-7
    A = np.eye(3 * n)
    rhs = np.zeros(3 * n)
    for e in range(n):
        K = mats[e, 4:13].reshape(3, 3)
        A[3 * e:3 * e + 3, :] -= K @ Qs[3 * e:3 * e + 3, :]
        rhs[3 * e:3 * e + 3] = K @ it.hext[e] + mats[e, 13:16]
    Mref = np.linalg.solve(A, rhs).reshape(n, 3)
    # FP64 system (oracle matrix): what the refined solve converges to
    A64 = np.eye(3 * n)
    for e in range(n):
        K = mats[e, 4:13].reshape(3, 3)
        A64[3 * e:3 * e + 3, :] -= K @ Qo[3 * e:3 * e + 3, :]
    M64 = np.linalg.solve(A64, rhs).reshape(n, 3)
    assert res[0] <= 1e-12
    print("refined vs FP64 system %.3e, stored system %.3e"
          % (np.abs(M - M64).max() / np.abs(M64).max(), np.abs(M - Mref).max() / np.abs(Mref).max()))
    assert np.abs(M - M64).max() <= 1e-11 * np.abs(M64).max()
    # without refinement the solve stops at the stored (FP32 off-diagonal) system
    M0 = np.stack([o.M for (o, _k, _s) in items])
    r2, M2, _H2 = it.matrix.relax(it.mats, it.hext, M0, 1e-12, 2000, refine=-1)
    assert r2.refined == 0 and r2.misfit <= 1e-12
    assert np.abs(M2 - Mref).max() <= 1e-11 * np.abs(Mref).max()


def test_relax_nonlinear_matches_fixed_point(oracle_off):
    """Nonlinear iron driven by a coil: the device solver reaches the FP64
    fixed point of the stored system (Newton reference), and agrees with the
    oracle's method-4 Gauss-Seidel (FP32 dot products, FP64-assembled matrix)
    to the conditioning of high-susceptibility iron."""
    print("case: relax nonlinear", flush=True)
    ids = models.tet_mesh((3, 3, 2), 4.0, jitter=0.1, seed=7, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    top = rad.ObjCnt([iron, src])
    P, items = models.packed(iron)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    k = rad.RlxPre(top)
    it = rad._reg[k]
    res = rad.RlxAuto(k, 1e-11, 3000)
    print("gpu relax", res, flush=True)
    M = np.stack([o.M for (o, _k, _s) in items])
    Qs = it.matrix.export()
    Qo, _jit = models.oracle_assemble(P)
    Mo, _H, iters, stats = OR.relax_m4(Qo.astype(np.float32), mats, it.hext, 1e-8, 20000)
    print("oracle m4", iters, stats, flush=True)
    # the refined solve reaches the FP64 fixed point of the oracle's matrix
    Mn = models.fp64_fixed_point(Qo, mats, it.hext, M0=Mo)
    scale = np.abs(Mn).max()
    assert res[0] <= 1e-11
    print("refined vs FP64 fixed point %.3e" % (np.abs(M - Mn).max() / scale), flush=True)
    # residual <= 1e-11 T times the conditioning of chi~2000 iron
    assert np.abs(M - Mn).max() <= 1e-7 * scale
    # FP32 dot products move the Gauss-Seidel fixed point by ~1.3e-4 here
    assert np.abs(M - Mo).max() <= 5e-4 * scale
    # unrefined: the fixed point of the stored matrix
    Ms = models.fp64_fixed_point(Qs, mats, it.hext, M0=Mo)
    r2, M2, _H2 = it.matrix.relax(it.mats, it.hext, np.zeros_like(M), 1e-11, 3000, refine=-1)
    assert r2.misfit <= 1e-11
    assert np.abs(M2 - Ms).max() <= 1e-7 * scale


def test_current_sources_match_oracle(oracle_off):
    print("case: currents", flush=True)
    a = rad.ObjArcCur([1, 2, 3], [5, 8], [0.2, 1.9], 3, 12, 2.5, "man", "z")
    b = rad.ObjRecCur([4, -2, 1], [2, 3, 4], [1.5, -2.0, 0.7])
    top = rad.ObjCnt([a, b])
    rad.TrfOrnt(top, rad.TrfRot([0, 0, 0], [1, 1, 0], 0.3))
    rng = np.random.default_rng(2)
    pts = rng.uniform(-12, 12, (300, 3))
    B = np.array(rad.Fld(top, "b", pts.tolist()))
    t = O.Trf.rotation([0, 0, 0], [1, 1, 0], 0.3)
    Bo = np.zeros_like(pts)
    for i, p in enumerate(pts):
        xl = np.linalg.solve(t.M, p - t.V)
        b1, _ = OR.arccur_BH(xl, [1, 2, 3], 5, 8, 0.2, 1.9, 3, 2.5, 12)
        b2, _ = OR.reccur_BH(xl, [4, -2, 1], [2, 3, 4], [1.5, -2.0, 0.7])
        Bo[i] = t.M @ (b1 + b2)
    assert np.abs(B - Bo).max() <= 1e-10 * np.abs(Bo).max()


def test_field_mesh_form(oracle_off):
    """A conforming tetrahedral mesh plus blocks, and a hexahedral ring: the
    mesh-form sweep (unique edges and faces) equals the per-element sweep and
    the oracle, inside and outside the elements, in FP64; FP32 stays within
    its per-element error."""
    print("case: field mesh form", flush=True)
    from radiacuda_amd import workloads as W
    rng = np.random.default_rng(41)
    ids = models.tet_mesh((4, 3, 2), 4.0, jitter=0.1, seed=42)
    ids.append(rad.ObjRecMag([24, 3, 2], [3, 4, 5], list(rng.uniform(-1, 1, 3))))
    ring, _p = W.halbach(n_seg=4, n_az=3, n_rad=2, n_ax=3, r_in=30.0, r_out=40.0, n_points=10)
    for top, box in ((rad.ObjCnt(ids), (-6, 30)), (ring, (-45, 45))):
        P, items = models.packed(top)
        M = rng.uniform(-1, 1, (len(items), 3))
        model = N.Model(P)
        ne, nf, ni = model.mesh_counts()
        print("mesh form: %d unique edges, %d unique faces, %d polyhedra" % (ne, nf, ni), flush=True)
        assert ni > 0 and model.field_form() == 2
        pts = rng.uniform(box[0], box[1], (3000, 3))
        Bm, Hm = model.field(M, pts)
        Bm32, _ = model.field(M, pts, precision=N.RH_MIXED)
        model.set_field_form(1)
        assert model.field_form() == 1
        Be, He = model.field(M, pts)
        Be32, _ = model.field(M, pts, precision=N.RH_MIXED)
        scale = np.abs(Be).max()
        assert np.abs(Bm - Be).max() <= 1e-12 * scale
        assert np.abs(Hm - He).max() <= 1e-12 * np.abs(He).max()
        e_mesh = np.abs(Bm32 - Bm).max() / scale
        e_elem = np.abs(Be32 - Be).max() / scale
        print("FP32 error: mesh form %.2e, per element %.2e" % (e_mesh, e_elem), flush=True)
        assert e_mesh <= max(4 * e_elem, 1e-6)
        sub = pts[:200]
        Bo, _Ho, _bad = models.oracle_field(P, M, sub)
        assert np.abs(Bm[:200] - Bo).max() <= 1e-10 * np.abs(Bo).max()
        model.close()


def _elem_geometry(o):
    """Vertices and 0-based faces of a polyhedron or a block."""
    if isinstance(o, O.RecMag):
        c = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
        return o.center + c * o.dims, [[v - 1 for v in f] for f in models.HEX_FACES]
    return o.verts, o.faces


def _near_points(items, d_rel):
    """Points d_rel element sizes off every face centroid (both sides, along
    the face normal) and off the midpoint of every edge of each element."""
    pts = []
    for (o, _k, _s) in items:
        V, faces = _elem_geometry(o)
        size = np.linalg.norm(V - V.mean(0), axis=1).max()
        for f in faces:
            c = V[f].mean(0)
            nrm = np.cross(V[f[1]] - V[f[0]], V[f[2]] - V[f[0]])
            nrm /= np.linalg.norm(nrm)
            pts += [c + d_rel * size * nrm, c - d_rel * size * nrm]
            for a, b in zip(f, f[1:] + f[:1]):
                u = np.cross(V[b] - V[a], [0.3, 0.7, 0.1])
                pts.append(0.5 * (V[a] + V[b]) + d_rel * size * u / np.linalg.norm(u))
    return np.array(pts)


@pytest.mark.parametrize("mirror", [False, True])
@pytest.mark.parametrize("form", [2, 1])
def test_fp32_field_near_elements(oracle_off, form, mirror):
    """The FP32 field sweep evaluates an element (or a mesh edge / face) in
    FP64 for every wave that holds a point within 1.5 element sizes of it, so
    the FP32 result stays within 2e-6 of the FP64 one (relative to the field
    scale) from 1e-1 down to 1e-6 element sizes off faces and edges, where
    FP32 cancellation alone loses every digit.  Tetrahedra (mesh triangles),
    hexahedra (mesh quads) and a block; mesh form and per-element form;
    with a mirror symmetry, also next to the mirrored copies."""
    print("case: fp32 near elements, form %d, mirror %s" % (form, mirror), flush=True)
    rng = np.random.default_rng(77)
    ids = models.tet_mesh((3, 2, 2), 4.0, jitter=0.1, seed=78)
    for c in ([16, 2, 3], [19, 2, 3], [16, 6, 3]):
        ids.append(models.hex_block(c, [3, 4, 6], rng.uniform(-1, 1, 3)))
    ids.append(rad.ObjRecMag([3, 13, 4], [4, 3, 5], list(rng.uniform(-1, 1, 3))))
    top = rad.ObjCnt(ids)
    if mirror:
        rad.TrfZerPara(top, [0, 0, -1], [0, 0, 1])  # copies in z' = -2 - z
    P, items = models.packed(top)
    M = rng.uniform(-1.5, 1.5, (len(items), 3))
    model = N.Model(P)
    if form == 1:
        model.set_field_form(1)
    assert model.field_form() == form
    for d_rel in (1e-1, 1e-3, 1e-5, 1e-6):
        pts = _near_points(items, d_rel)
        if mirror:
            pts = np.concatenate([pts, pts * [1, 1, -1] - [0, 0, 2]])
        B64, H64 = model.field(M, pts, precision=N.RH_FP64)
        B32, H32 = model.field(M, pts, precision=N.RH_MIXED)
        eb = np.abs(B32 - B64).max() / np.abs(B64).max()
        eh = np.abs(H32 - H64).max() / np.abs(H64).max()
        print("distance %.0e sizes, %d points: FP32 error B %.2e, H %.2e" % (d_rel, len(pts), eb, eh),
              flush=True)
        assert eb <= 2e-6 and eh <= 2e-6
    model.close()


@pytest.mark.parametrize("t", [1.0, 1e-2, 1e-4, 1e-6, 1e-8])
def test_sliver_tetrahedron_on_device(oracle_off, t):
    """The reference's flatness sweep (tests/test_sliver_tet_asm.py) on the
    device: assembly (FP64, FP32 storage) and fields finite down to 1e-8 mm,
    the FP64 self block and H within 1e-10 of 50-digit values, FP32 within the
    reference's 5e-5; the drop-in solve runs on the GPU backend and stays
    finite."""
    from tests.test_hostcheck import SLIVER_OBS, sliver_model, sliver_mp50
    print("case: sliver t=%g" % t, flush=True)
    top = sliver_model(t)
    P, items = models.packed(top)
    H50, Q50 = sliver_mp50(t, P["M"][0], P["M"][1], SLIVER_OBS)
    model = N.Model(P)
    for prec in (N.RH_FP64, N.RH_MIXED):
        Q = model.assemble(prec)
        Qs = Q.export()
        Q.close()
        assert np.isfinite(Qs).all()
        tol = 1e-10 if prec == N.RH_FP64 else 1e-6
        assert np.abs(Qs[0:3, 0:3] - Q50).max() <= tol
    for prec, tol in ((N.RH_FP64, 1e-10), (N.RH_MIXED, 5e-5)):
        B, H = model.field(P["M"], SLIVER_OBS, precision=prec)
        assert np.isfinite(B).all() and np.isfinite(H).all()
        assert np.abs(H - H50).max() <= tol * np.abs(H50).max()
    model.close()
    r = rad.Solve(top, 1e-5, 1000)
    assert rad.UtiAsmLastBackend() in ("gpu", "gpu_streaming")
    Bs = np.array([rad.Fld(top, "b", list(p)) for p in SLIVER_OBS])
    assert r[0] <= 1e-5 and np.isfinite(Bs).all()


def test_repeated_solve_and_field_are_bit_identical(oracle_off):
    """Every reduction runs in a fixed order: the same solve and the same
    field sweep twice give identical bits (the reference repeats a GPU solve
    and compares %.17e strings, tests/test_gpu_asm_cpu_solve.py:187-200)."""
    print("case: repeated solve and field", flush=True)
    ids = models.tet_mesh((4, 3, 2), 4.0, jitter=0.1, seed=11, M=[0, 0, 0])
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm([2118, 1.362], [63.06, 0.2605], [17.138, 0.4917]))
    src = rad.ObjRecCur([-8, 6, 4], [4, 12, 8], [0, 0, 2.0])
    top = rad.ObjCnt([iron, src])
    k = rad.RlxPre(top)
    it = rad._reg[k]
    n = len(it.items)
    pts = np.random.default_rng(4).uniform(-6, 20, (5000, 3))
    runs = []
    for _ in range(2):
        r, M, H = it.matrix.relax(it.mats, it.hext, np.zeros((n, 3)), 1e-10, 2000)
        B32 = it.model.field(M, pts, precision=N.RH_MIXED)[0]
        B64 = it.model.field(M, pts, precision=N.RH_FP64)[0]
        runs.append((r.iters, M, H, B32, B64))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1:], runs[1][1:]):
        assert np.array_equal(a, b)
    rad.UtiDel(k)


@pytest.mark.parametrize("shape", ["recmag", "rotated_polyhedron"])
@pytest.mark.parametrize("mr", [1.2, [0.4, -0.7, 0.9]])
def test_matlin_anisotropic_cube_closed_form(shape, mr):
    """One cube of MatLin([ksi_par, ksi_perp], mr) in a uniform background
    field H0.  A cube's self field at its centre is -M/3 in any orientation,
    so the relaxed magnetization solves (I + K/3) M = K H0 + M_r exactly, with
    K = ksi_perp I + (ksi_par - ksi_perp) L L^T built here from the
    reference's component formula (radmater.h:292-303): L is the element's
    initial magnetization direction for a scalar mr (radmater.h:235-250), the
    direction of the vector mr otherwise, and M_r = |mr| L.  B at the centre
    is H0 + 2M/3."""
    print("case: MatLin cube closed form", shape, mr, flush=True)
    kpar, kper = 0.06, 0.17
    m0 = np.array([0.3, 0.5, 0.8])
    if shape == "recmag":
        c = rad.ObjRecMag([1.0, -2.0, 3.0], [10.0, 10.0, 10.0], list(m0))
        ctr = np.array([1.0, -2.0, 3.0])
    else:
        R = np.linalg.qr(np.random.default_rng(9).normal(size=(3, 3)))[0]
        ctr = np.array([4.0, 1.0, -2.0])
        c = models.hex_block(ctr, [10.0, 10.0, 10.0], m0, rot=R)
    rad.MatApl(c, rad.MatLin([kpar, kper], mr))
    H0 = np.array([0.05, -0.2, 0.3])
    top = rad.ObjCnt([c, rad.ObjBckg(list(H0))])
    k = rad.RlxPre(top)
    res = rad.RlxAuto(k, 1e-13, 1000)
    assert res[0] <= 1e-13
    M = np.stack([o.M for (o, _k, _s) in rad._reg[k].items])[0]
    if np.ndim(mr) == 0:
        L = m0 / np.linalg.norm(m0)
        Mr = mr * L
    else:
        Mr = np.asarray(mr, dtype=np.float64)
        L = Mr / np.linalg.norm(Mr)
    d = kpar - kper
    K = np.array([[kpar * L[0] ** 2 + kper * (L[1] ** 2 + L[2] ** 2), d * L[0] * L[1], d * L[0] * L[2]],
                  [d * L[0] * L[1], kpar * L[1] ** 2 + kper * (L[0] ** 2 + L[2] ** 2), d * L[1] * L[2]],
                  [d * L[0] * L[2], d * L[1] * L[2], kpar * L[2] ** 2 + kper * (L[0] ** 2 + L[1] ** 2)]])
    Mx = np.linalg.solve(np.eye(3) + K / 3.0, K @ H0 + Mr)
    assert np.abs(M - Mx).max() <= 1e-10 * np.abs(Mx).max()
    B = np.array(rad.Fld(top, "b", list(ctr)))
    Bx = H0 + 2.0 * Mx / 3.0
    assert np.abs(B - Bx).max() <= 1e-10 * np.abs(Bx).max()


def test_sweep_results_depend_on_point_order_only_at_rounding_level():
    """The wave-uniform atan2 tiers (rh_atan2_wave) pick a polynomial per wave,
    so a point's result depends on which points share its wave, within the
    tiers' agreement: ~1e-17 of an angle in FP64, a few 1e-8 in FP32.  A
    permuted and a strided subset of the same points give the same fields to
    1e-12 (FP64) and 1e-6 (FP32 integrals) of the field scale."""
    print("case: point order / batching of the mesh-form sweep", flush=True)
    ids = models.tet_mesh((4, 3, 3), 5.0, jitter=0.1, seed=12)
    top = rad.ObjCnt(ids)
    P, _items = models.packed(top)
    rng = np.random.default_rng(13)
    pts = np.concatenate([rng.uniform(-5, 25, (30000, 3)), rng.uniform(-60, 80, (30000, 3))])
    perm = rng.permutation(len(pts))
    with N.Model(P) as m:
        assert m.field_form() == 2
        for prec, tol in ((N.RH_FP64, 1e-12), (N.RH_MIXED, 1e-6)):
            B = m.field(P["M"], pts, precision=prec)[0]
            Bp = np.empty_like(B)
            Bp[perm] = m.field(P["M"], pts[perm], precision=prec)[0]
            Bs = m.field(P["M"], pts[::7], precision=prec)[0]
            s = np.abs(B).max()
            dp, ds = np.abs(Bp - B).max() / s, np.abs(Bs - B[::7]).max() / s
            print("prec %d: permuted %.2e, subset %.2e" % (prec, dp, ds), flush=True)
            assert dp <= tol and ds <= tol
