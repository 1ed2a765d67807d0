"""Seeded model builders shared by the tests (Radia API calls only)."""
import numpy as np

from radiacuda_amd import rad
from radiacuda_amd import objects as O

TET_FACES = [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]]
# Kuhn split of a unit cube (corner index = 4i + 2j + k) into 6 tetrahedra
KUHN = [[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]]
HEX_FACES = [[1, 2, 4, 3], [5, 6, 8, 7], [1, 2, 6, 5], [3, 4, 8, 7], [1, 3, 7, 5], [2, 4, 8, 6]]


def tet_mesh(shape, h, origin=(0, 0, 0), jitter=0.1, seed=0, M=None):
    """Structured hexahedral grid, interior vertices jittered by +-jitter*h,
    each cell split into 6 tetrahedra.  Returns the list of object ids."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    g = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij"), -1).astype(float) * h + np.asarray(origin, float)
    inner = np.zeros(g.shape[:3], bool)
    inner[1:-1, 1:-1, 1:-1] = True
    g[inner] += rng.uniform(-jitter * h, jitter * h, (inner.sum(), 3))
    ids = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = np.array([g[i + a, j + b, k + d] for a in (0, 1) for b in (0, 1) for d in (0, 1)])
                for t in KUHN:
                    m = rng.uniform(-1, 1, 3) if M is None else M
                    ids.append(rad.ObjPolyhdr(c[t].tolist(), TET_FACES, list(m)))
    return ids


def hex_block(center, dims, M, rot=None):
    c = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)]) * dims
    if rot is not None:
        c = c @ np.asarray(rot).T
    c = c + np.asarray(center, float)
    return rad.ObjPolyhdr(c.tolist(), HEX_FACES, list(M))


def mixed_model(seed=1, sym=True):
    """Tets + blocks + a hexahedral polyhedron, with symmetries and a rotation."""
    rng = np.random.default_rng(seed)
    ids = tet_mesh((2, 1, 1), 4.0, origin=(1, 2, 3), jitter=0.08, seed=seed)
    ids.append(rad.ObjRecMag([3, 11, 4], [2, 3, 4], list(rng.uniform(-1, 1, 3))))
    ids.append(rad.ObjRecMag([-5, 4, 6], [3, 2, 2], list(rng.uniform(-1, 1, 3))))
    ids.append(hex_block([-6, 1, 2], [3, 3, 2], rng.uniform(-1, 1, 3)))
    a = rad.ObjCnt(ids[:6])
    b = rad.ObjCnt(ids[6:])
    if sym:
        rad.TrfZerPara(a, [0, 0, 0], [0, 0, 1])
        rad.TrfOrnt(b, rad.TrfRot([1, 1, 1], [0.2, 0.3, 1.0], 0.4))
        rad.TrfZerPerp(b, [0, -1, 0], [0, 1, 0])
    return rad.ObjCnt([a, b])


def packed(obj, relax_only=False):
    """Flat element arrays of every magnetizable element under obj."""
    items = [it for it in O.walk(rad._reg, obj)
             if isinstance(it[0], (O.RecMag, O.Polyhedron)) and not getattr(it[0], "is_current", False)]
    return O.pack_elements(items), items


def oracle_assemble(P):
    """Oracle matrix with randomization off (FldLenRndSw('off')).  Blocks whose
    collocation point lies exactly on a face plane of the source element are
    0/0 in that mode (Radia's own behaviour); those blocks are recomputed with
    Radia's default length randomization (1e-9 mm jitter).  Returns (Q, mask)
    with mask True for the blocks that needed the jitter."""
    from oracle import oracle as OR
    OR.set_randomization(False)
    om = OR.prepare_model(P)
    Q = OR.assemble(om)
    n = len(P["type"])
    bad = ~np.isfinite(Q).reshape(n, 3, n, 3).all(axis=(1, 3))
    if bad.any():
        OR.set_randomization(True)
        Qr = OR.assemble(OR.prepare_model(P))
        OR.set_randomization(False)
        for i, j in np.argwhere(bad):
            Q[3 * i:3 * i + 3, 3 * j:3 * j + 3] = Qr[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    return Q, bad


def block_rel(A, B, n, mask=None, tol_masked=None):
    """Max over 3x3 blocks of |A-B|/max|B| (per block); blocks in mask are
    returned separately."""
    d = np.abs(A - B).reshape(n, 3, n, 3).max(axis=(1, 3))
    s = np.abs(B).reshape(n, 3, n, 3).max(axis=(1, 3))
    r = d / np.maximum(s, 1e-300)
    if mask is None:
        return r.max()
    return (r[~mask].max() if (~mask).any() else 0.0), (r[mask].max() if mask.any() else 0.0)


def oracle_field(P, M, pts):
    """Oracle B, H with randomization off; points where that is 0/0 (exactly on
    a face plane) are recomputed with Radia's default length randomization.
    Returns (B, H, mask of recomputed points)."""
    from oracle import oracle as OR
    OR.set_randomization(False)
    om = OR.prepare_model(P)
    B, H = OR.field(om, M, pts)
    bad = ~(np.isfinite(B).all(axis=1) & np.isfinite(H).all(axis=1))
    if bad.any():
        OR.set_randomization(True)
        Br, Hr = OR.field(OR.prepare_model(P), M, np.asarray(pts)[bad])
        OR.set_randomization(False)
        B[bad] = Br
        H[bad] = Hr
    return B, H, bad


def _tab_f_and_df(t, h):
    """|M|(h) and d|M|/dh of a node table t [L, 3] (cubic Hermite between
    nodes, constant below the first, last slope beyond the last)."""
    x, y, d = t[:, 0], t[:, 1], t[:, 2]
    L = len(x)
    i = int(np.searchsorted(x, h, side="right")) - 1
    if i < 0:
        return float(y[0]), 0.0
    if i >= L - 1:
        return float(y[-1] + (h - x[-1]) * d[-1]), float(d[-1])
    s = x[i + 1] - x[i]
    u = h - x[i]
    sl = (y[i + 1] - y[i]) / s
    c2 = (3 * sl - 2 * d[i] - d[i + 1]) / s
    c3 = (-2 * sl + d[i] + d[i + 1]) / (s * s)
    return float(y[i] + u * (d[i] + u * (c2 + u * c3))), float(d[i] + u * (2 * c2 + 3 * u * c3))


def mat_M_and_D(rec, H, tab=None):
    """M(H) and dM/dH of a material record (kinds 0 linear iso, 1 linear
    aniso, 2 tanh-sum formula, 3 tabulated law in tab), FP64, for the
    test-side reference solver."""
    kind = int(rec[0])
    mr = rec[13:16]
    if kind == 0:
        return rec[1] * H + mr, rec[1] * np.eye(3)
    if kind == 1:
        K = rec[4:13].reshape(3, 3)
        return K @ H + mr, K
    h = float(np.linalg.norm(H))
    if kind == 3:
        t = np.asarray(tab, dtype=np.float64)[3 * int(rec[1]):3 * int(rec[1] + rec[2])].reshape(-1, 3)
        if h == 0:
            return mr.copy(), t[0, 2] * np.eye(3)
        f, fp = _tab_f_and_df(t, h)
        u = H / h
        chi = f / h
        return chi * H + mr, chi * np.eye(3) + (fp - chi) * np.outer(u, u)
    ms, ks = rec[2:2 + int(rec[1])], rec[5:5 + int(rec[1])]
    f = sum(m * np.tanh(k * h / m) for m, k in zip(ms, ks) if m != 0)
    def sech2(x):
        e = np.exp(-2.0 * abs(x))
        return 4.0 * e / (1.0 + e) ** 2
    fp = sum(k * sech2(k * h / m) for m, k in zip(ms, ks) if m != 0)
    if h == 0:
        return mr.copy(), fp * np.eye(3)
    u = H / h
    chi = f / h
    return chi * H + mr, chi * np.eye(3) + (fp - chi) * np.outer(u, u)


def fp64_fixed_point(Q, mats, hext, M0=None, tol=1e-14, max_iter=100, tab=None):
    """Newton polish of M = M_mat(Q M + hext) in FP64 with the exact Jacobian
    (dense, small n), started from M0 (e.g. the oracle's Gauss-Seidel result):
    the reference the device solver is checked against on the very matrix the
    device stores."""
    n = len(mats)
    if M0 is None:
        M = np.array([r[13:16] for r in mats], dtype=np.float64).ravel()
    else:
        M = np.asarray(M0, dtype=np.float64).ravel().copy()
    for _ in range(max_iter):
        H = (Q @ M).reshape(n, 3) + hext
        F = np.zeros(3 * n)
        D = np.zeros((3 * n, 3 * n))
        for e in range(n):
            m, d = mat_M_and_D(mats[e], H[e], tab)
            F[3 * e:3 * e + 3] = M[3 * e:3 * e + 3] - m
            D[3 * e:3 * e + 3, 3 * e:3 * e + 3] = d
        if np.abs(F).max() <= tol:
            break
        J = np.eye(3 * n) - D @ Q
        M = M - np.linalg.solve(J, F)
    return M.reshape(n, 3)


# Steel M(H) curve of the reference's hybrid undulator example
# (examples/RADIA_Example03.py:59-69): H in A/m, M in tesla.
STEEL_H_AM = [0.8, 1.5, 2.2, 3.6, 5, 6.8, 9.8, 18, 28, 37.5, 42, 55, 71.5, 80, 85, 88, 92, 100, 120,
              150, 200, 300, 400, 600, 800, 1000, 2000, 4000, 6000, 10000, 25000, 40000]
STEEL_M_T = [0.000998995, 0.00199812, 0.00299724, 0.00499548, 0.00699372, 0.00999145, 0.0149877,
             0.0299774, 0.0499648, 0.0799529, 0.0999472, 0.199931, 0.49991, 0.799899, 0.999893,
             1.09989, 1.19988, 1.29987, 1.41985, 1.49981, 1.59975, 1.72962, 1.7995, 1.89925,
             1.96899, 1.99874, 2.09749, 2.19497, 2.24246, 2.27743, 2.28958, 2.28973]
MU0 = 4e-7 * np.pi


def steel_points():
    """(|H| in tesla, |M| in tesla) pairs of the example's steel."""
    return [[h * MU0, m] for h, m in zip(STEEL_H_AM, STEEL_M_T)]


def hybrid_undulator(n_periods=2, period=46.0, gap=20.0, offset=1.0, pole=(45.0, 5.0, 25.0),
                     magnet_xz=(65.0, 45.0), pole_div=(2, 2, 5), magnet_div=(1, 3, 1)):
    """Hybrid undulator of the reference's third example: steel poles with the
    tabulated curve above, NdFeB blocks (MatLin([0.06, 0.17], 1.2)) magnetized
    along -+y in turn; the quarter x > 0, z < 0 is built and three symmetry
    planes complete it.  Pieces along y from the mid-plane: a half pole, then
    n_periods of (magnet, pole), then a half magnet.  Returns (top, steel)."""
    steel = rad.MatSatIsoTab(steel_points())
    ndfeb = rad.MatLin([0.06, 0.17], 1.2)
    t_mag = 0.5 * period - pole[1]
    seq = [("pole", 0.5 * pole[1], 0)]
    sign = -1
    for _ in range(n_periods):
        seq += [("magnet", t_mag, sign), ("pole", pole[1], 0)]
        sign = -sign
    seq.append(("magnet", 0.5 * t_mag, sign))
    top = rad.ObjCnt([])
    y = 0.0
    for kind, t, s in seq:
        yc = y + 0.5 * t
        y += t
        if kind == "pole":
            b = rad.ObjRecMag([0.25 * pole[0], yc, -0.5 * (pole[2] + gap)], [0.5 * pole[0], t, pole[2]],
                              [0, 0, 0])
            rad.ObjDivMag(b, list(pole_div))
            rad.MatApl(b, steel)
        else:
            b = rad.ObjRecMag([0.25 * magnet_xz[0], yc, -0.5 * (magnet_xz[1] + gap) - offset],
                              [0.5 * magnet_xz[0], t, magnet_xz[1]], [0, s, 0])
            rad.ObjDivMag(b, list(magnet_div))
            rad.MatApl(b, ndfeb)
        rad.ObjAddToCnt(top, [b])
    rad.TrfZerPerp(top, [0, 0, 0], [1, 0, 0])
    rad.TrfZerPara(top, [0, 0, 0], [0, 0, 1])
    rad.TrfZerPerp(top, [0, 0, 0], [0, 1, 0])
    return top, steel


def tilted_minus_z(items, tol=1e-7):
    """Indices of polyhedra with a face whose outward unit normal n satisfies
    0 < |1 + n_z| <= tol (tol = the reference's RelLengthTol) without being
    exactly -z.  For such a face the reference's face rotation
    (radvlpgn.cpp:357-371) takes the axis-aligned frame diag(1, -1, -1) and
    drops the face's tilt, so the oracle, which restates it, is off by up to
    ~1e-3 of the element's field there; a face exactly normal to -z loses
    nothing."""
    out = []
    for e, it in enumerate(items):
        o = it[0]
        if not hasattr(o, "faces"):
            continue
        V = np.asarray(o.verts, dtype=np.float64)
        c = V.mean(axis=0)
        for f in o.faces:
            a, b, d = V[f[0]], V[f[1]], V[f[2]]
            nv = np.cross(b - a, d - a)
            nv /= np.linalg.norm(nv)
            if np.dot(nv, a - c) < 0:
                nv = -nv
            if abs(1.0 + nv[2]) <= tol and (nv[0] != 0.0 or nv[1] != 0.0):
                out.append(e)
                break
    return np.asarray(out, dtype=np.int64)
