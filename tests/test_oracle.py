"""Pins the CPU oracle (no GPU): known answers, closed forms, invariants,
brute-force quadrature and the reference's own golden solve results."""
import json
import os

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as OR
from radiacuda_amd import rad
from tests import golden_models, models

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(autouse=True)
def _rand_off():
    OR.set_randomization(False)
    yield
    OR.set_randomization(False)


CUBE = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
CUBE_FACES = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]


def test_example01_known_answer():
    """examples/RADIA_Example01.py:16-21: 1 mm cube, M = [-0.5, 1, 0.7] T,
    B at [0.52, 0.6, 0.7] close to [0.12737, 0.028644, 0.077505] (5 digits)."""
    B, _ = OR.recmag_BH([0.52, 0.6, 0.7], [0, 0, 0], [1, 1, 1], [-0.5, 1, 0.7])
    assert np.allclose(B, [0.12737, 0.028644, 0.077505], rtol=0, atol=6e-6)


def test_block_closed_form_invariants():
    # cube at its own centre: Q = -I/3 exactly (demagnetizing factors 1/3)
    assert np.allclose(OR.recmag_Q([0, 0, 0], [0, 0, 0], [1, 1, 1]), -np.eye(3) / 3, atol=1e-15)
    rng = np.random.default_rng(0)
    D = np.array([1.0, 2.0, 3.0])
    for _ in range(30):
        p = rng.uniform(-0.45, 0.45, 3) * D
        Q = OR.recmag_Q(p, [0, 0, 0], D)
        assert abs(np.trace(Q) + 1.0) < 1e-13      # trace -1 inside
        assert np.abs(Q - Q.T).max() < 1e-15       # Hessian: symmetric
        p = rng.uniform(-4, 4, 3) * D
        if np.all(np.abs(p) <= D / 2):
            continue
        Q = OR.recmag_Q(p, [0, 0, 0], D)
        assert abs(np.trace(Q)) < 1e-13            # trace 0 outside


def test_block_far_field_is_a_dipole():
    D = np.array([1.0, 2.0, 1.5])
    V = np.prod(D)
    r = np.array([40.0, -25.0, 33.0]) * 3
    R = np.linalg.norm(r)
    rh = r / R
    Qd = V * (3 * np.outer(rh, rh) - np.eye(3)) / (4 * np.pi * R**3)
    Q = OR.recmag_Q(r, [0, 0, 0], D)
    assert np.abs(Q - Qd).max() < 1e-4 * np.abs(Qd).max()


def test_polyhedron_cube_equals_block():
    p = OR.Poly(CUBE * [2, 3, 1] + [1, 2, 3], CUBE_FACES)
    rng = np.random.default_rng(1)
    M = [0.3, -0.7, 1.1]
    for x in rng.uniform(-4, 6, (40, 3)):
        Bp, Hp = p.BH(x, M)
        Br, Hr = OR.recmag_BH(x, [1, 2, 3], [2, 3, 1], M)
        assert np.abs(Bp - Br).max() <= 1e-13 * max(1.0, np.abs(Br).max())
    assert np.allclose(p.centroid, [1, 2, 3])


def _face_charge_field(verts, faces, M, x):
    """H = (1/4pi) sum_f sigma_f int_f (x - y)/|x - y|^3 dA by adaptive quadrature
    over triangle faces (sigma = M.n, outward n)."""
    verts = np.asarray(verts, float)
    c = verts.mean(0)
    H = np.zeros(3)
    for f in faces:
        a, b, d = verts[f[0]], verts[f[1]], verts[f[2]]
        n = np.cross(b - a, d - a)
        area2 = np.linalg.norm(n)
        n /= area2
        if np.dot(a - c, n) < 0:
            n = -n
        sig = np.dot(M, n)
        for k in range(3):
            def g(v, u, k=k):
                y = a + u * (b - a) + v * (d - a)
                r = x - y
                return r[k] / np.linalg.norm(r) ** 3
            val, _ = integrate.dblquad(g, 0, 1, 0, lambda u: 1 - u, epsabs=1e-13, epsrel=1e-11)
            H[k] += sig * val * area2 / (4 * np.pi)
    return H


def test_tetrahedron_matches_quadrature():
    verts = [[0.1, 0.0, 0.2], [2.0, 0.3, 0.1], [0.4, 1.8, 0.0], [0.6, 0.5, 1.7]]
    faces = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
    p = OR.Poly(verts, faces)
    M = np.array([0.4, -0.2, 0.9])
    for x in ([3.0, 2.5, -1.0], [-1.5, 0.7, 2.2]):
        _, H = p.BH(x, M)
        Hq = _face_charge_field(verts, faces, M, np.array(x))
        assert np.abs(H - Hq).max() <= 1e-8 * np.abs(Hq).max()


def test_polyhedron_trace_and_symmetry():
    verts = [[0.1, 0.0, 0.2], [2.0, 0.3, 0.1], [0.4, 1.8, 0.0], [0.6, 0.5, 1.7]]
    p = OR.Poly(verts, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])
    Q = p.Q(p.centroid)
    assert abs(np.trace(Q) + 1) < 1e-13
    Q = p.Q([4.0, -1.0, 2.0])
    assert abs(np.trace(Q)) < 1e-13 and np.abs(Q - Q.T).max() < 1e-13


def test_block_subdivision_superposition():
    M = [0.2, 0.5, -0.8]
    x = [2.3, -1.1, 0.7]
    B, _ = OR.recmag_BH(x, [0, 0, 0], [2, 2, 2], M)
    acc = np.zeros(3)
    for c in CUBE:
        b, _ = OR.recmag_BH(x, c, [1, 1, 1], M)
        acc += b
    assert np.abs(acc - B).max() < 1e-14


def _biot_savart_box(x, c, d, J):
    """B = 1e-4 * integral J x (x - y)/|x - y|^3 dV (tplquad)."""
    c, d = np.asarray(c, float), np.asarray(d, float)
    G = np.zeros(3)
    for k in range(3):
        def g(z, y, xx, k=k):
            r = np.asarray(x) - np.array([xx, y, z])
            return r[k] / np.linalg.norm(r) ** 3
        G[k], _ = integrate.tplquad(g, c[0] - d[0] / 2, c[0] + d[0] / 2, c[1] - d[1] / 2,
                                    c[1] + d[1] / 2, c[2] - d[2] / 2, c[2] + d[2] / 2,
                                    epsabs=1e-12, epsrel=1e-10)
    return 1e-4 * np.cross(J, G)


def test_block_current_matches_biot_savart():
    J = np.array([1.5, -2.0, 0.7])
    x = [3.0, 2.0, -2.5]
    B, _ = OR.reccur_BH(x, [0, 0, 0], [2, 3, 1], J)
    Bq = _biot_savart_box(x, [0, 0, 0], [2, 3, 1], J)
    assert np.abs(B - Bq).max() <= 1e-7 * np.abs(Bq).max()


def test_arc_current_converges_to_closed_loop_axis_field():
    """A full annulus of small section: on-axis Bz -> mu0 I R^2/(2 (R^2+z^2)^1.5)."""
    R, w, h, J = 50.0, 0.2, 0.2, 10.0
    I = J * w * h  # A
    z = 20.0
    B, _ = OR.arccur_BH([0, 0, z], [0, 0, 0], R - w / 2, R + w / 2, 0.0, 2 * np.pi, h, J, 720)
    mu0_4pi = 1e-7  # T m / A ; lengths in mm -> multiply by 1e3
    Bz = 2 * np.pi * mu0_4pi * I * (R * 1e-3) ** 2 / ((R * 1e-3) ** 2 + (z * 1e-3) ** 2) ** 1.5
    assert abs(B[2] - Bz) <= 2e-4 * abs(Bz)
    assert abs(B[0]) < 1e-6 * abs(Bz) and abs(B[1]) < 1e-6 * abs(Bz)


def test_materials_closed_forms():
    rec = np.zeros(16)
    rec[0] = 2
    rec[1] = 3
    rec[2:5] = [1.362, 0.2605, 0.4917]
    rec[5:8] = [2118, 63.06, 17.138]
    H = np.array([0.01, -0.02, 0.005])
    h = np.linalg.norm(H)
    f = sum(ms * np.tanh(ks * h / ms) for ms, ks in zip(rec[2:5], rec[5:8]))
    assert np.allclose(OR.mat_M(rec, H), f * H / h, rtol=1e-15)


def test_relax_m4_linear_reaches_direct_solution():
    ids = models.tet_mesh((2, 2, 1), 4.0, jitter=0.1, seed=2, M=[0, 0, 1.0])
    top = rad.ObjCnt(ids)
    P, items = models.packed(top)
    n = len(items)
    om = OR.prepare_model(P)
    Q = OR.assemble(om)
    ksi = 5.0
    mats = np.zeros((n, 16))
    mats[:, 0] = 0
    mats[:, 1] = ksi
    hext = np.tile([0.0, 0.02, 0.1], (n, 1))
    Qf = Q.astype(np.float32)
    # the method-4 sweep forms its dot products in FP32 (gmvectf.h:210-216),
    # which floors the attainable misfit near 1e-9 T
    M, _H, it, stats = OR.relax_m4(Qf, mats, hext, 1e-8, 5000)
    A = np.eye(3 * n) - ksi * Qf.astype(np.float64)
    Mref = np.linalg.solve(A, ksi * hext.ravel()).reshape(n, 3)
    assert stats[0] <= 1e-8
    assert np.abs(M - Mref).max() <= 1e-6 * np.abs(Mref).max()


@pytest.mark.parametrize("kind", ["small", "mixed", "large"])
def test_oracle_pipeline_matches_reference_golden(kind):
    """tests/_mpi_solve_reference.json of the reference (cpu_asm_cpu_relax):
    60 random tetrahedra with two symmetries (with 12 blocks: "mixed"), or
    1,100 random RecMag blocks without symmetries ("large"), MatSatIsoFrm
    iron and a racetrack coil; B at two points within 1e-6 of the field scale
    (the reference test's own bound, tests/test_mpi_solve.py:116-123).  The
    "large" solve stops at its 3,000-iteration cap with misfit 0.174 (chi =
    20,000 and overlapping blocks), so matching it pins the method-4 sweep's
    whole trajectory, FP32 dot products included (about 25 s)."""
    OR.set_randomization(True)  # the reference model keeps Radia's default randomization
    gold = json.load(open(os.path.join(HERE, "golden", "mpi_solve_reference.json")))[kind]
    ref = gold["cpu_asm_cpu_relax"]
    B, misfit = golden_models.oracle_solve(kind)
    scale = np.abs(np.array(ref["b"])).max()
    err = np.abs(B - np.array(ref["b"])).max() / scale
    assert err < 1e-6, err
    assert abs(misfit - ref["misfit"]) <= 1e-2 * ref["misfit"]
