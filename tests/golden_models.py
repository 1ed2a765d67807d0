"""The models behind the reference's golden solve results
(tests/_mpi_solve_reference.json), rebuilt through the drop-in API exactly as
tests/test_mpi_solve.py:33-80 builds them ("small", "mixed" and "large"), and
the oracle's end-to-end solve of them (assembly, method-4 relaxation, field,
coil)."""
import numpy as np

from oracle import oracle as OR
from radiacuda_amd import objects as O
from radiacuda_amd import rad

POINTS = [[25.0, 10.0, 0.0], [60.0, 30.0, 5.0]]


def build(kind="small"):
    rad.UtiDelAll()
    rng = np.random.default_rng(11)
    if kind == "large":
        # 1,100 RecMag blocks (tests/test_mpi_solve.py:37-46), no symmetries
        ids = []
        for _ in range(1100):
            center = rng.uniform([-80, -80, -40], [80, 80, 40]).tolist()
            dims = rng.uniform(2.0, 4.0, 3).tolist()
            ids.append(rad.ObjRecMag(center, dims))
        iron = rad.ObjCnt(ids)
        rad.MatApl(iron, rad.MatSatIsoFrm([20000, 2], [0.1, 2], [0.1, 2]))
        coil = rad.ObjRaceTrk([0, 0, 30], [70, 90], [0, 0], 15, 10, 3.0, "man", "z")
        return rad.ObjCnt([iron, coil]), iron, coil
    faces = [[1, 2, 3], [1, 4, 2], [2, 4, 3], [3, 4, 1]]
    ids = []
    for _ in range(60):
        base = rng.uniform([20, 5, 5], [90, 40, 40])
        verts = base + rng.uniform(2, 10, size=(4, 3))
        ids.append(rad.ObjPolyhdr(verts.tolist(), faces))
    if kind == "mixed":
        for i in range(3):
            for j in range(2):
                for k in range(2):
                    ids.append(rad.ObjRecMag([30.0 + i * 12, 10.0 + j * 11, 8.0 + k * 9],
                                             [10.0, 8.0, 7.0]))
    iron = rad.ObjCnt(ids)
    rad.TrfZerPerp(iron, [0, 0, 0], [1, 0, 0])
    rad.TrfZerPara(iron, [0, 0, 0], [0, 0, 1])
    mat = rad.MatSatIsoFrm([20000, 2], [0.1, 2], [0.1, 2])
    rad.MatApl(iron, mat)
    coil = rad.ObjRaceTrk([0, 0, 30], [70, 90], [0, 0], 15, 10, 3.0, "man", "z")
    return rad.ObjCnt([iron, coil]), iron, coil


def coil_field_oracle(coil, pts):
    B = np.zeros((len(pts), 3))
    for o, _k, stack in O.walk(rad._reg, coil):
        for t in O._expand(stack):
            for i, p in enumerate(pts):
                xl = np.linalg.solve(t.M, np.asarray(p) - t.V)
                if isinstance(o, O.ArcCur):
                    b, _ = OR.arccur_BH(xl, o.center, o.rmin, o.rmax, o.phimin, o.phimax,
                                        o.height, o.j, o.nseg)
                else:
                    b, _ = OR.reccur_BH(xl, o.center, o.dims, o.J)
                B[i] += t.s * (t.M @ b)
    return B


def oracle_solve(kind="small", prec=1e-5, max_iter=3000, return_state=False):
    top, iron, coil = build(kind)
    items = [it for it in O.walk(rad._reg, iron)]
    P = O.pack_elements(items)
    om = OR.prepare_model(P)
    X = OR.colloc(om)
    Hlab = coil_field_oracle(coil, X)
    n = len(items)
    hext = np.zeros((n, 3))
    for e in range(n):
        c = P["copy_off"][e]
        hext[e] = P["cS"][c] * np.linalg.solve(P["cM"][c], Hlab[e])
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    Q = OR.assemble(om)
    M, _H, _it, stats = OR.relax_m4(Q.astype(np.float32), mats, hext, prec, max_iter)
    B, _ = OR.field(om, M.ravel(), POINTS)
    B = B + coil_field_oracle(coil, POINTS)
    if return_state:
        return B, stats[0], M
    return B, stats[0]
