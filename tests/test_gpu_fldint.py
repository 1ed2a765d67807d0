"""Field integrals (FldInt) on the GPU field, pinned by what the mathematics
fixes: H = -grad(phi) with phi continuous, so the integral of H along any
infinite line is 0, and that of B along a line through a uniformly magnetized
body is M . (chord vector); a finite integral against SciPy quadrature of the
oracle's field; an infinite integral against finite integrals over growing
segments (the tails fall off as 1/L^2)."""
import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as OR
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh():
    rad.UtiDelAll()
    OR.set_randomization(False)
    yield
    rad.UtiDelAll()


def test_infinite_integrals_of_a_uniform_block_and_tetrahedra():
    print("case: FldInt inf, block + tetrahedra", flush=True)
    M = np.array([0.3, -0.2, 1.1])
    b = rad.ObjRecMag([1, 2, 3], [10, 12, 14], M.tolist())
    t = rad.ObjPolyhdr([[20, 0, 0], [30, 0, 0], [20, 10, 0], [20, 0, 10]],
                       [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]], [0.0, 0.5, 0.4])
    rad.ObjDivMag(t, [2, 2, 2])
    top = rad.ObjCnt([b, t])
    # a line along z through the block only (chord 14 along z)
    p1, p2 = [2.0, 3.5, -5.0], [2.0, 3.5, 5.0]
    ibz = rad.FldInt(top, "inf", "ibz", p1, p2)
    ihz = rad.FldInt(top, "inf", "ihz", p1, p2)
    print("inf ibz %.15g (M_z * 14 = %.15g), ihz %.3e" % (ibz, M[2] * 14, ihz), flush=True)
    assert abs(ibz - M[2] * 14.0) <= 1e-9 * abs(M[2] * 14.0)
    assert abs(ihz) <= 1e-9 * abs(M[2] * 14.0)
    # a slanted line through block and tetrahedra: the longitudinal integral of
    # H vanishes, that of B is the sum of M . chord over the bodies crossed
    d = np.array([1.0, 0.1, 0.05])
    d /= np.linalg.norm(d)
    q1 = np.array([0.0, 3.0, 2.0])
    ib = np.array(rad.FldInt(top, "inf", "ib", q1.tolist(), (q1 + d).tolist()))
    ih = np.array(rad.FldInt(top, "inf", "ih", q1.tolist(), (q1 + d).tolist()))
    assert abs(ih @ d) <= 1e-9 * np.abs(ib).max()
    # chord through the block: the parameter range inside all six slabs
    lo, hi = -np.inf, np.inf
    c, h = np.array([1.0, 2.0, 3.0]), np.array([5.0, 6.0, 7.0])
    for a in range(3):
        t1, t2 = (c[a] - h[a] - q1[a]) / d[a], (c[a] + h[a] - q1[a]) / d[a]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    chord_block = max(0.0, hi - lo)
    # the tetrahedron pieces: the chord inside x + y + z <= 30 etc. by sampling
    s = np.linspace(-100, 200, 3000001)
    X = q1[None, :] + s[:, None] * d[None, :]
    inside_t = (X[:, 0] >= 20) & (X[:, 1] >= 0) & (X[:, 2] >= 0) & (X.sum(1) <= 30.0)
    chord_tet = inside_t.sum() * (s[1] - s[0])
    want = M @ d * chord_block + np.array([0.0, 0.5, 0.4]) @ d * chord_tet
    print("inf ib.d %.12g, M.chord %.12g (tet chord %.4f)" % (ib @ d, want, chord_tet), flush=True)
    assert abs(ib @ d - want) <= 1e-3 * abs(want)  # the sampled tetrahedron chord limits this check
    # exact when the line misses the tetrahedra
    q2 = np.array([0.0, 3.0, 2.0])
    e = np.array([0.2, 1.0, 0.1]) / np.linalg.norm([0.2, 1.0, 0.1])
    ib2 = np.array(rad.FldInt(top, "inf", "ib", q2.tolist(), (q2 + e).tolist()))
    lo, hi = -np.inf, np.inf
    for a in range(3):
        t1, t2 = (c[a] - h[a] - q2[a]) / e[a], (c[a] + h[a] - q2[a]) / e[a]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    assert abs(ib2 @ e - M @ e * (hi - lo)) <= 1e-9 * abs(M @ e * (hi - lo))


def test_finite_integral_against_quadrature_of_the_oracle_field():
    print("case: FldInt fin vs scipy quad of the oracle", flush=True)
    ids = models.tet_mesh((2, 2, 1), 5.0, jitter=0.1, seed=3, M=[0.2, -0.4, 0.9])
    top = rad.ObjCnt(ids)
    P, items = models.packed(top)
    M = np.stack([o.M for (o, _k, _s) in items])
    om = OR.prepare_model(P)
    p1, p2 = np.array([-3.0, 4.0, 2.5]), np.array([14.0, 6.0, 2.0])
    u = (p2 - p1) / np.linalg.norm(p2 - p1)
    Lseg = np.linalg.norm(p2 - p1)

    def fx(tt, k):
        B, _H = OR.field(om, M, (p1 + tt * u)[None, :])
        return B[0, k]

    brk = [0.0] + rad._line_breaks(top, p1, u)
    brk = sorted({min(max(t, 0.0), Lseg) for t in brk} | {Lseg})
    ref = np.zeros(3)
    for k in range(3):
        for a, b in zip(brk[:-1], brk[1:]):
            if b > a:
                ref[k] += integrate.quad(fx, a, b, args=(k,), epsabs=1e-14, epsrel=1e-12, limit=200)[0]
    got = np.array(rad.FldInt(top, "fin", "ib", p1.tolist(), p2.tolist()))
    print("fin ib", got, "quad", ref, flush=True)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_infinite_integral_is_the_limit_of_finite_ones():
    print("case: FldInt inf as the limit of fin", flush=True)
    ids = models.tet_mesh((2, 1, 1), 6.0, jitter=0.1, seed=5, M=[0.1, 0.3, -0.8])
    top = rad.ObjCnt(ids)
    p1, p2 = [-2.0, 3.0, 9.0], [10.0, 3.5, 9.5]  # passes above the mesh
    inf = np.array(rad.FldInt(top, "inf", "ib", p1, p2))
    u = np.subtract(p2, p1) / np.linalg.norm(np.subtract(p2, p1))
    c = np.array(p1)
    for L, tol in ((1e3, 1e-4), (1e5, 1e-8)):
        fin = np.array(rad.FldInt(top, "fin", "ib", (c - L * u).tolist(), (c + L * u).tolist()))
        print("L %.0e: fin - inf %.3e" % (L, np.abs(fin - inf).max() / np.abs(inf).max()), flush=True)
        assert np.abs(fin - inf).max() <= tol * np.abs(inf).max()
