"""Drop-in subset of Radia's Python API (``import radiacuda_amd.rad as rad``).

Function names, argument order, defaults and return shapes follow the
reference's Python module (cpp/src/clients/python/radpy.cpp, method table at
:3341-3434).  Geometry, materials and transforms live in the host object model
(objects.py); assembly, relaxation and field evaluation run in the HIP
library through its C ABI (include/radhip.h) -- there is no CPU fallback.
"""
import math

import numpy as np

from . import objects as O
from . import _native as N

_reg = {}
_next = [1]
_gen = [0]             # structure generation (invalidates cached device models)
_rand = ["on"]         # FldLenRndSw state


def _bump():
    _gen[0] += 1


def _add(o):
    k = _next[0]
    _next[0] += 1
    _reg[k] = o
    _bump()
    return k


def _obj(k):
    try:
        return _reg[int(k)]
    except (KeyError, TypeError, ValueError):
        raise RuntimeError("Radia::Error003: no object with index %r" % (k,))


def _vec3(v, name="vector"):
    a = np.asarray(v, dtype=np.float64).ravel()
    if a.size != 3:
        raise RuntimeError("%s must have 3 components" % name)
    return a


# ------------------------------------------------------------------ objects
def ObjRecMag(P, L, M=(0, 0, 0)):
    """Rectangular block with centre P, dimensions L, magnetization M (radpy.cpp:144)."""
    return _add(O.RecMag(_vec3(P), _vec3(L), _vec3(M)))


def ObjRecCur(P, L, J):
    """Current-carrying rectangular block (radpy ObjRecCur)."""
    return _add(O.RecMag(_vec3(P), _vec3(L), (0, 0, 0), _vec3(J)))


def ObjPolyhdr(V, F, M=(0, 0, 0), *args, **kw):
    """Convex polyhedron from vertices V and 1-based faces F (radpy.cpp:230-357)."""
    faces = [[int(i) - 1 for i in f] for f in F]
    return _add(O.Polyhedron(np.asarray(V, dtype=np.float64), faces, _vec3(M)))


_ERROR063 = ("Radia::Error063::::Multiple extruded polygon can not be generated from this input: "
             "incorrect definition of horizontal slice polygons.")
_ERROR111 = ("Radia::Error111::::Multiple extruded polygon can not be generated from this input: "
             "Non-convex horizontal slice polygon encountered.")


def ObjMltExtPgn(slices, M=(0, 0, 0)):
    """Convex polyhedra between consecutive horizontal layers
    [[[x, y], ...], z] (radpy.cpp ObjMltExtPgn; radapl1.cpp:859-880,
    radapl4.cpp:435-510, 1036-1064).  A layer holds 1 point (an apex), 2 points
    (a ridge) or a convex polygon, and consecutive layers may hold different
    numbers of points; the z values are distinct and strictly monotonic.  Each
    layer pair becomes one polyhedron, the convex hull of the two layers
    (objects.layer_segment); several are grouped in a container.  Errors as
    the reference's: Error111 for a non-convex layer polygon, Error063 for
    any other improper layer structure."""
    M = _vec3(M)
    try:
        layers = [(np.asarray(p, dtype=np.float64).reshape(-1, 2), float(z)) for p, z in slices]
    except (TypeError, ValueError):
        raise RuntimeError(_ERROR063)
    if len(layers) < 2 or max(len(p) for p, _z in layers) < 2:
        raise RuntimeError(_ERROR063)
    dz = np.diff([z for _p, z in layers])
    if not ((dz > 0).all() or (dz < 0).all()):
        raise RuntimeError(_ERROR063)
    ids = []
    for (pa, za), (pb, zb) in zip(layers[:-1], layers[1:]):
        try:
            verts, faces = O.layer_segment(pa, za, pb, zb)
        except O.LayerError as e:
            raise RuntimeError(_ERROR111 if e.code == 111 else _ERROR063)
        ids.append(_add(O.Polyhedron(verts, faces, M)))
    return ids[0] if len(ids) == 1 else ObjCnt(ids)


def ObjArcCur(P, R, Phi, h, nseg, j, man_auto="man", a="z"):
    """Arc current of rectangular section (radapl1.cpp:1664-1710)."""
    if str(man_auto).lower() != "man":
        raise NotImplementedError("only the 'man' (fixed-segment) arc integration is supported")
    o = O.ArcCur(_vec3(P), R[0], R[1], Phi[0], Phi[1], h, nseg, j)
    k = _add(o)
    _orient_along(k, P, a)
    return k


def ObjRaceTrk(P, R, L, h, nseg, j, man_auto="man", a="z"):
    """Racetrack coil: four 90-degree arcs plus straight blocks (radapl1.cpp:1879-1970)."""
    P = _vec3(P)
    rmin, rmax = float(R[0]), float(R[1])
    lx, ly = float(L[0]), float(L[1])
    pi = 3.14159265358979324
    arcs = []
    spec = [((P[0] + 0.5 * lx, P[1] + 0.5 * ly), (0.0, 0.5 * pi)),
            ((P[0] - 0.5 * lx, P[1] + 0.5 * ly), (0.5 * pi, pi)),
            ((P[0] - 0.5 * lx, P[1] - 0.5 * ly), (pi, 1.5 * pi)),
            ((P[0] + 0.5 * lx, P[1] - 0.5 * ly), (1.5 * pi, 1.999999999999 * pi))]
    for (cx, cy), (p0, p1) in spec:
        arcs.append(_add(O.ArcCur((cx, cy, P[2]), rmin, rmax, p0, p1, h, nseg, j)))
    members = list(arcs)
    if ly != 0.0:
        eff = 0.5 * (lx + rmin + rmax)
        d = (rmax - rmin, ly, h)
        members.append(ObjRecCur((P[0] + eff, P[1], P[2]), d, (0.0, j, 0.0)))
        members.append(ObjRecCur((P[0] - eff, P[1], P[2]), d, (0.0, -j, 0.0)))
    if lx != 0.0:
        eff = 0.5 * (ly + rmin + rmax)
        d = (lx, rmax - rmin, h)
        members.append(ObjRecCur((P[0], P[1] + eff, P[2]), d, (-j, 0.0, 0.0)))
        members.append(ObjRecCur((P[0], P[1] - eff, P[2]), d, (j, 0.0, 0.0)))
    k = ObjCnt(members)
    _orient_along(k, P, a)
    return k


def _orient_along(k, P, a):
    a = str(a).lower()
    if a == "z":
        return
    # rotation taking the local z axis to the requested axis about P
    if a == "x":
        t = O.Trf.rotation(P, (0, 1, 0), 0.5 * math.pi)
    elif a == "y":
        t = O.Trf.rotation(P, (1, 0, 0), -0.5 * math.pi)
    else:
        raise RuntimeError("orientation must be 'x', 'y' or 'z'")
    _obj(k).add_trf(t, 1)
    _bump()


def ObjBckg(B):
    return _add(O.Background(_vec3(B)))


def ObjCnt(members):
    for k in members:
        _obj(k)
    return _add(O.Container([int(k) for k in members]))


def ObjAddToCnt(cnt, members):
    c = _obj(cnt)
    if not isinstance(c, O.Container):
        raise RuntimeError("Radia::Error004: not a container")
    c.members.extend(int(k) for k in members)
    _bump()
    return cnt


def ObjCntStuf(obj):
    o = _obj(obj)
    return list(o.members) if isinstance(o, O.Container) else [obj]


def ObjCntSize(obj):
    o = _obj(obj)
    return len(o.members) if isinstance(o, O.Container) else 0


def ObjDivMag(obj, k, *args):
    """Subdivide a block into kx*ky*kz blocks, or a polyhedron by planes normal
    to the axes of its own frame ([k1,k2,k3] or [[k1,q1],...]); the object
    becomes a container of the pieces (radrec.cpp:971-1080 for blocks;
    radvlpgn.cpp:2176-2300 for polyhedra: the planes of each direction are
    placed over the extent of the whole polyhedron, the pieces ordered by x
    slice, then y, then z, empty pieces dropped)."""
    o = _obj(obj)
    if isinstance(o, O.Container):
        for m in list(o.members):
            ObjDivMag(m, k, *args)
        return obj
    spec = []
    for kk in k:
        if isinstance(kk, (list, tuple)):
            spec.append((int(round(kk[0])), float(kk[1])))
        else:
            spec.append((int(round(kk)), 1.0))
    if any(cnt < 1 for cnt, _q in spec):
        raise RuntimeError("subdivision numbers must be >= 1")
    if isinstance(o, O.Polyhedron):
        return _div_polyhedron(obj, o, spec)
    if not isinstance(o, O.RecMag):
        raise NotImplementedError("ObjDivMag is supported for blocks and polyhedra")
    edges = []
    for ax in range(3):
        cnt, q = spec[ax]
        L = o.dims[ax]
        if cnt < 1:
            raise RuntimeError("subdivision numbers must be >= 1")
        if cnt == 1 or q == 1.0:
            w = np.full(cnt, L / cnt)
        else:
            r = q ** (1.0 / (cnt - 1))
            w = r ** np.arange(cnt)
            w = w * (L / w.sum())
        e = o.center[ax] - 0.5 * L + np.concatenate([[0.0], np.cumsum(w)])
        edges.append(e)
    kids = []
    for ix in range(len(edges[0]) - 1):
        for iy in range(len(edges[1]) - 1):
            for iz in range(len(edges[2]) - 1):
                lo = np.array([edges[0][ix], edges[1][iy], edges[2][iz]])
                hi = np.array([edges[0][ix + 1], edges[1][iy + 1], edges[2][iz + 1]])
                b = O.RecMag(0.5 * (lo + hi), hi - lo, o.M, o.J)
                b.mat = o.mat
                kids.append(_add(b))
    c = O.Container(kids)
    c.trfs = o.trfs
    _reg[int(obj)] = c
    _bump()
    return obj


def ObjFullMag(P, L, M, K, grp, mat, rgb=(0, 0, 0)):
    """Block with centre P, dimensions L and magnetization M, subdivided by K,
    given material mat (0: none) and added to container grp (radpy.cpp:612)."""
    k = ObjRecMag(P, L, M)
    ObjDivMag(k, K)
    if int(mat) != 0:
        MatApl(k, mat)
    g = _obj(grp)
    if not isinstance(g, O.Container):
        raise RuntimeError("Radia::Error016: not a container")
    g.members.append(int(k))
    _bump()
    return k


def ObjMltExtRtg(slices, M=(0, 0, 0)):
    """Polyhedra between consecutive rectangles [[centre(3), [wx, wy]], ...]
    in planes normal to z (radpy.cpp:428-500): ObjMltExtPgn with the
    rectangles as slices."""
    if len(slices) < 2:
        raise RuntimeError("Incorrect argument: ObjMltExtRtg, number of slices / layers should be >= 2")
    pg = []
    for (c, d) in slices:
        c = _vec3(c)
        hx, hy = 0.5 * float(d[0]), 0.5 * float(d[1])
        pg.append([[[c[0] - hx, c[1] - hy], [c[0] + hx, c[1] - hy], [c[0] + hx, c[1] + hy],
                    [c[0] - hx, c[1] + hy]], c[2]])
    return ObjMltExtPgn(pg, M)


def ObjDegFre(obj):
    """Degrees of freedom of the relaxation: 3 per magnetizable element
    (radpy.cpp:1396)."""
    n = 0
    for o, _k, _s in O.walk(_reg, int(obj)):
        if isinstance(o, (O.RecMag, O.Polyhedron)) and not getattr(o, "is_current", False):
            n += 3
    return n


def ObjDrwAtr(obj, rgb=(0, 0, 0), thcn=0.001):
    """Drawing attributes: graphics are not part of this build, so this only
    checks the object and returns it."""
    _obj(obj)
    return obj


def ObjCutMag(obj, P, N, opt=""):
    """Cuts obj by the plane through P with normal N into new objects on its
    lower and upper side; returns their keys (the object itself when the
    plane misses it) and leaves obj as it was (radapl3.cpp:1347-1420).  In
    a container, members wholly on one side join that side's container as
    they are.  The
    plane is taken in the frame of the object's own coordinates
    ('Frame->Loc'); blocks are cut as polyhedra."""
    if opt and "loc" not in str(opt).lower():
        raise NotImplementedError("ObjCutMag: only 'Frame->Loc' is supported")
    nrm = np.asarray(_vec3(N), dtype=np.float64)
    if not np.linalg.norm(nrm) > 0.0:
        raise RuntimeError("Incorrect argument: ObjCutMag")
    nrm = nrm / np.linalg.norm(nrm)
    off = float(nrm @ np.asarray(_vec3(P)))
    lo, hi = _cut(int(obj), nrm, off)
    if lo is None or hi is None:
        return [int(obj)]
    return [lo, hi]


def _as_polyhedron(o):
    if isinstance(o, O.Polyhedron):
        return o.verts, o.faces
    c, h = np.asarray(o.center), 0.5 * np.asarray(o.dims)
    V = np.array([[c[0] + sx * h[0], c[1] + sy * h[1], c[2] + sz * h[2]]
                  for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    return V, HEX_FACES0


HEX_FACES0 = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]


def _cut(key, nrm, off):
    o = _obj(key)
    if isinstance(o, O.Container):
        los, his = [], []
        for m in o.members:
            a, b = _cut(int(m), nrm, off)
            if a is not None:
                los.append(a)
            if b is not None:
                his.append(b)
        if not his:
            return key, None  # wholly on one side: the container itself
        if not los:
            return None, key
        out = []
        for part in (los, his):
            c = O.Container(part)
            c.trfs = list(o.trfs)
            out.append(_add(c))
        return out[0], out[1]
    if not isinstance(o, (O.RecMag, O.Polyhedron)) or getattr(o, "is_current", False):
        raise NotImplementedError("ObjCutMag: only magnetized blocks and polyhedra are cut")
    V, F = _as_polyhedron(o)
    size = float(np.abs(V.max(0) - V.min(0)).max())
    below, above = O.cut_convex(V, F, nrm, off, 1e-10 * size)
    if above is None:
        return key, None  # wholly on one side: the object itself
    if below is None:
        return None, key
    keys = []
    for piece in (below, above):
        if piece is None:
            keys.append(None)
            continue
        b = O.Polyhedron(piece[0], piece[1], o.M)
        b.mat = o.mat
        b.trfs = list(o.trfs)
        keys.append(_add(b))
    _bump()
    return keys[0], keys[1]


def _div_polyhedron(obj, o, spec):
    V = o.verts
    size = float(np.abs(V.max(0) - V.min(0)).max())
    tol = 1e-10 * size
    pieces = [(V, [list(f) for f in o.faces])]
    for ax in range(3):
        cnt, q = spec[ax]
        if cnt <= 1:
            continue
        nrm = np.zeros(3)
        nrm[ax] = 1.0
        planes = O.division_planes(float(V[:, ax].min()), float(V[:, ax].max()), cnt, q)
        nxt = []
        for pv, pf in pieces:
            rest = (pv, pf)
            for c in planes:
                below, rest = O.cut_convex(rest[0], rest[1], nrm, c, tol)
                if below is not None:
                    nxt.append(below)
                if rest is None:
                    break
            if rest is not None:
                nxt.append(rest)
        pieces = nxt
    kids = []
    for pv, pf in pieces:
        b = O.Polyhedron(pv, pf, o.M)
        b.mat = o.mat
        kids.append(_add(b))
    c = O.Container(kids)
    c.trfs = o.trfs
    _reg[int(obj)] = c
    _bump()
    return obj


def ObjThckPgn(xc, lx, pgn, a="x", M=(0, 0, 0)):
    """Prism of thickness lx centred at xc along axis a over the polygon pgn,
    whose 2D points map to (y, z) for a = 'x', (z, x) for 'y' and (x, y) for
    'z' (radpy.cpp:179-226, radapl1.cpp:318-470).  An axis-aligned rectangle
    becomes a block, as in the reference; other polygons must be convex."""
    if isinstance(a, (list, tuple, np.ndarray)):
        a, M = "x", a
    a = str(a).lower()
    if a not in ("x", "y", "z") or float(lx) <= 0.0:
        raise RuntimeError("Incorrect argument: ObjThckPgn")
    P = np.asarray(pgn, dtype=np.float64).reshape(-1, 2)
    if len(P) < 3:
        raise RuntimeError("Incorrect argument: ObjThckPgn, incorrect polygon definition")
    ax = "xyz".index(a)
    other = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[ax]
    M = _vec3(M)
    lo, hi = float(xc) - 0.5 * float(lx), float(xc) + 0.5 * float(lx)
    xs, ys = np.unique(P[:, 0]), np.unique(P[:, 1])
    if len(P) == 4 and len(xs) == 2 and len(ys) == 2 and \
            all(((p[0] == xs).any() and (p[1] == ys).any()) for p in P) and len({tuple(p) for p in P}) == 4:
        ctr = np.zeros(3)
        dims = np.zeros(3)
        ctr[ax], dims[ax] = float(xc), float(lx)
        ctr[other[0]], dims[other[0]] = 0.5 * (xs[0] + xs[1]), xs[1] - xs[0]
        ctr[other[1]], dims[other[1]] = 0.5 * (ys[0] + ys[1]), ys[1] - ys[0]
        return _add(O.RecMag(ctr, dims, M))
    # convexity: every turn of the loop has the same sign
    d = np.roll(P, -1, 0) - P
    turn = d[:, 0] * np.roll(d, -1, 0)[:, 1] - d[:, 1] * np.roll(d, -1, 0)[:, 0]
    if not ((turn > 0).all() or (turn < 0).all()):
        raise NotImplementedError("ObjThckPgn: only convex polygons are supported")
    n = len(P)
    verts = np.zeros((2 * n, 3))
    for k, (u, v) in enumerate(P):
        for layer, t in ((0, lo), (1, hi)):
            w = verts[layer * n + k]
            w[ax], w[other[0]], w[other[1]] = t, u, v
    faces = [list(range(n)), list(range(n, 2 * n))]
    for i in range(n):
        j = (i + 1) % n
        faces.append([i, j, n + j, n + i])
    return _add(O.Polyhedron(verts, faces, M))


def ObjM(obj):
    """[[centre, M], ...] of the magnetizable elements (radpy ObjM)."""
    out = []
    for o, _k, _s in O.walk(_reg, int(obj)):
        if isinstance(o, O.RecMag):
            out.append([list(o.center), list(o.M)])
        elif isinstance(o, O.Polyhedron):
            out.append([list(o.verts.mean(0)), list(o.M)])
    return out[0] if len(out) == 1 else out


def ObjSetM(obj, M):
    M = _vec3(M)
    for o, _k, _s in O.walk(_reg, int(obj)):
        if isinstance(o, (O.RecMag, O.Polyhedron)) and not getattr(o, "is_current", False):
            o.M = M.copy()
    return obj


# ------------------------------------------------------------------ transforms
def TrfTrsl(V):
    return _add(O.Trf.translation(_vec3(V)))


def TrfRot(P, V, phi):
    return _add(O.Trf.rotation(_vec3(P), _vec3(V), float(phi)))


def TrfPlSym(P, N):
    return _add(O.Trf.plane_symmetry(_vec3(P), _vec3(N)))


def TrfInv():
    return _add(O.Trf.field_inversion())


def TrfCmbL(orig, trf):
    t = _obj(orig)
    _reg[int(orig)] = _obj(trf).compose(t)
    return orig


def TrfCmbR(orig, trf):
    t = _obj(orig)
    _reg[int(orig)] = t.compose(_obj(trf))
    return orig


def _trf(k):
    t = _obj(k)
    if not isinstance(t, O.Trf):
        raise RuntimeError("Radia::Error012: not a transformation")
    return t


def TrfOrnt(obj, trf):
    _obj(obj).add_trf(_trf(trf), 1)
    _bump()
    return obj


def TrfMlt(obj, trf, mlt):
    _obj(obj).add_trf(_trf(trf), int(mlt))
    _bump()
    return obj


def TrfZerPerp(obj, P, N):
    """Plane symmetry, multiplicity 2 (radentry.cpp:1802-1840)."""
    _obj(obj).add_trf(O.Trf.plane_symmetry(_vec3(P), _vec3(N)), 2)
    _bump()
    return obj


def TrfZerPara(obj, P, N):
    """Plane symmetry combined with field inversion, multiplicity 2 (radentry.cpp:1802-1840)."""
    t = O.Trf.plane_symmetry(_vec3(P), _vec3(N))
    t.s = -1.0
    _obj(obj).add_trf(t, 2)
    _bump()
    return obj


# ------------------------------------------------------------------ materials
def MatLin(ksi, mr):
    """MatLin([ksipar, ksiper], mr | [mrx, mry, mrz]) (radmater.h:164-303)."""
    if np.isscalar(ksi):
        kpar = kper = float(ksi)
    else:
        kpar, kper = float(ksi[0]), float(ksi[1])
    if np.isscalar(mr):
        return _add(O.Material(O.MAT_LIN_ANISO, kpar=kpar, kper=kper, mr=float(mr)))
    return _add(O.Material(O.MAT_LIN_ANISO, kpar=kpar, kper=kper, mr_vec=_vec3(mr)))


def MatSatIsoFrm(*pairs):
    """MatSatIsoFrm([ksi1,ms1],[ksi2,ms2],[ksi3,ms3]) (radentry.cpp:792-813)."""
    ks = [float(p[0]) for p in pairs]
    ms = [float(p[1]) for p in pairs]
    return _add(O.Material(O.MAT_NONLIN_FRM, ks=ks, ms=ms))


def MatSatIsoTab(points):
    """MatSatIsoTab([[H1, M1], [H2, M2], ...]): isotropic law |M|(|H|) from a
    table, H and M in tesla (radentry.cpp:817-846, radmater.h:436-451)."""
    pts = [(float(p[0]), float(p[1])) for p in points]
    if pts[0][0] < 0 or pts[0][1] < 0 or any(b[0] < a[0] for a, b in zip(pts, pts[1:])):
        raise RuntimeError("Radia::Error071: H values must be non-negative and increasing")
    x, y, d = O.tab_nodes(pts)
    return _add(O.Material(O.MAT_NONLIN_TAB, x=x, y=y, d=d))


# Radia's standard materials (radmater.cpp:714-731): linear anisotropic
# (ksi_par, ksi_perp, Mr) and tanh-sum saturating laws (ms_1..3, ksi_1..3).
_STD_MATERIALS = {
    "NDFEB": ("lin", (0.06, 0.17, 1.2)),
    "SMCO5": ("lin", (0.005, 0.04, 0.85)),
    "SM2CO17": ("lin", (0.005, 0.04, 1.05)),
    "FERRITE": ("lin", (0.07, 0.2, 0.35)),
    "XC06": ("sat", ((1.362, 0.2605, 0.4917), (2118.0, 63.06, 17.138))),
    "STEEL37": ("sat", ((1.1488, 0.4268, 0.4759), (1596.3, 133.11, 18.713))),
    "STEEL42": ("sat", ((1.441, 0.2912, 0.3316), (968.66, 24.65, 8.3))),
    "AFK502": ("sat", ((1.788, 0.437, 0.115), (10485.0, 241.5, 7.43))),
    "AFK1": ("sat", ((1.704, 0.493, 0.152), (2001.0, 38.56, 1.24))),
}


def MatStd(name, mr=0.0):
    """Standard material by name (case-insensitive); for the permanent
    magnets mr > 0 replaces the remanence (radpy.cpp:1900, radmater.cpp:733-770)."""
    ent = _STD_MATERIALS.get(str(name).upper())
    if ent is None:
        raise RuntimeError("Radia::Error008: unknown material name")
    kind, par = ent
    if kind == "lin":
        kpar, kper, br = par
        return MatLin([kpar, kper], float(mr) if float(mr) > 0.0 else br)
    ms, ks = par
    return MatSatIsoFrm(*[[ks[i], ms[i]] for i in range(3)])


def MatApl(obj, mat):
    m = _obj(mat)
    if not isinstance(m, O.Material):
        raise RuntimeError("Radia::Error007: not a material")
    for o, _k, _s in O.walk(_reg, int(obj)):
        if isinstance(o, (O.RecMag, O.Polyhedron)) and not getattr(o, "is_current", False):
            o.mat = m
    _bump()
    return obj


def MatMvsH(obj, comp, H):
    o = None
    for o, _k, _s in O.walk(_reg, int(obj)):
        if getattr(o, "mat", None) is not None:
            break
    if o is None or o.mat is None:
        raise RuntimeError("object has no material")
    rec, tab = O.pack_materials([(o, None, None)])
    M = N.mat_M(rec[0], _vec3(H), tab)
    return _component(M, comp or "m")


# ------------------------------------------------------------------ relaxation
class _Interaction:
    pass


def _ext_field_at(ext_items, pts):
    """H of the external sources at lab points (element, currents, background)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    H = np.zeros_like(pts)
    mags = [it for it in ext_items if isinstance(it[0], (O.RecMag, O.Polyhedron))
            and not getattr(it[0], "is_current", False)]
    curs = [it for it in ext_items if isinstance(it[0], (O.ArcCur,)) or
            (isinstance(it[0], O.RecMag) and it[0].is_current)]
    if mags:
        packed = O.pack_elements(mags)
        with N.Model(packed) as m:
            _b, h = m.field(packed["M"], pts, want_b=False)
        H += h
    if curs:
        H += N.current_field(O.pack_currents(curs), pts)
    for o, _k, stack in ext_items:
        if isinstance(o, O.Background):
            for t in O._expand(stack):
                H += t.s * (t.M @ o.B)
    return H


def RlxPre(obj, srcobj=0, use_gpu=True):
    """Builds the interaction structure of obj (radentry.cpp:923-930): device
    model, device-resident interaction matrix and the external field."""
    rel, ext = O.flatten(_reg, int(obj))
    if srcobj:
        _r2, e2 = O.flatten(_reg, int(srcobj))
        ext = ext + [(o, k, s) for (o, k, s) in _r2 + e2]
    if not rel:
        raise RuntimeError("Radia::Error011: no relaxable elements (apply a material)")
    it = _Interaction()
    it.items = rel
    it.packed = O.pack_elements(rel)
    it.model = N.Model(it.packed)
    it.matrix = _assemble_with_policy(it.model)
    X = it.model.colloc()
    Hlab = _ext_field_at(ext, X) if ext else np.zeros_like(X)
    # into element frames: s * Minv of the main copy (radintrc.cpp:1231)
    co = it.packed["copy_off"]
    hext = np.zeros_like(Hlab)
    for e in range(len(rel)):
        c = co[e]
        hext[e] = it.packed["cS"][c] * np.linalg.solve(it.packed["cM"][c], Hlab[e])
    it.hext = hext
    it.mats, it.tab = O.pack_materials(rel)
    it.H = np.zeros_like(hext)
    return _add(it)


_ERROR602 = ("Radia::Error602::::GPU interaction-matrix assembly could not be completed "
             "(e.g. insufficient GPU memory) and the GPU fallback policy is 'break'. Reduce the "
             "model size, or set UtiGpuFallback('gpu_streaming') to stream the matrix rows that "
             "do not fit from host memory.")


def _assemble_with_policy(model):
    """Device assembly under the UtiGpuFallback policy (radintrc.cpp:796-805):
    'break' refuses a matrix beyond the device budget; 'gpu_streaming' and
    'cpu' keep the rows that fit resident and stream the rest from pinned
    host memory (there is no CPU assembly here: the device streams instead).
    RADHIP_MATRIX_DEVICE_BYTES and RADHIP_MATRIX_RESIDENT_FRACTION override
    the budget and the split (as the reference's RADGPU_FORCE_STREAM hooks)."""
    import os
    budget = int(os.environ.get("RADHIP_MATRIX_DEVICE_BYTES", "0"))
    frac = float(os.environ.get("RADHIP_MATRIX_RESIDENT_FRACTION", "-1"))
    try:
        return model.assemble(N.RH_MIXED, device_bytes=budget, resident_fraction=frac,
                              allow_streaming=True if frac >= 0 else "policy")
    except N.RadHipError as e:
        if "exceeds the device budget" in str(e) or "does not fit" in str(e):
            raise RuntimeError(_ERROR602) from e
        raise


def _opt(opts, name, default):
    for s in opts:
        for part in str(s).split(","):
            if "->" in part:
                k, v = part.split("->", 1)
                if k.strip().lower() == name.lower():
                    return v.strip()
    return default


def RlxAuto(intrc, prec, maxiter, meth=4, *opts):
    """Relaxes to prec (T) within maxiter iterations; returns
    [misfit, max|M|, max|H|, iterations] (radentry.h:746-761).  meth 11 runs
    the device Newton-Krylov solver (misfit = RMS |M_mat(H) - M|, iterations
    = matrix passes, maxiter caps them, as radGPU_RelaxNK reports,
    radgpurlx.h:78-86); every other method number runs the device
    Anderson-accelerated block Jacobi.  Both work on the FP32-stored matrix,
    then one FP64 refinement pass (option 'Refine->k' sets k passes,
    'Refine->-1' none)."""
    it = _obj(intrc)
    if not isinstance(it, _Interaction):
        raise RuntimeError("Radia::Error014: not an interaction structure")
    zero_m = _opt(opts, "ZeroM", "True").lower() not in ("false", "0", "no")
    omega = float(_opt(opts, "omega", "0"))
    refine = int(_opt(opts, "Refine", "0"))
    M0 = np.stack([o.M for (o, _k, _s) in it.items])
    res, M, H = it.matrix.relax(it.mats, it.hext, M0, prec, maxiter, keep_m=not zero_m,
                                omega0=omega, refine=refine, tab=it.tab,
                                method=11 if int(meth) == 11 else 9)
    for e, (o, _k, _s) in enumerate(it.items):
        o.M = M[e].copy()
    it.H = H
    it.last = res
    return [res.misfit, res.max_m, res.max_h, res.iters]


def RlxMan(intrc, meth, iters, rlx=0.0):
    """Runs iters iterations of the relaxation from the current
    magnetization and returns [misfit, max|M|, max|H|]
    (radpy.cpp:2271-2296).  Every method number runs the device
    Anderson-accelerated block Jacobi (its first step under-relaxed by rlx
    when 0 < rlx <= 1), so the state after a given number of iterations is
    not the reference's method-by-method trajectory; the fixed point is the
    same.  The iterations work on the stored matrix without the FP64
    refinement pass."""
    if int(iters) < 1:
        raise RuntimeError("Incorrect argument: RlxMan")
    w = float(rlx)
    opts = ("ZeroM->False", "Refine->-1") + (("omega->%r" % w,) if 0.0 < w <= 1.0 else ())
    r = RlxAuto(intrc, 0.0, int(iters), 4, *opts)
    return r[:3]


def Solve(obj, prec, maxiter, meth=4):
    """RlxPre + RlxAuto (radentry.cpp:1842-1855)."""
    k = RlxPre(obj)
    try:
        return RlxAuto(k, prec, maxiter, meth)
    finally:
        UtiDel(k)


# ------------------------------------------------------------------ field
_cache = {}


def _field_model(obj):
    key = int(obj)
    ent = _cache.get(key)
    if ent is not None and ent[0] == _gen[0]:
        return ent[1]
    items = list(O.walk(_reg, key))
    mags = [it for it in items if isinstance(it[0], (O.RecMag, O.Polyhedron))
            and not getattr(it[0], "is_current", False)]
    curs = [it for it in items if isinstance(it[0], O.ArcCur) or
            (isinstance(it[0], O.RecMag) and it[0].is_current)]
    bck = [it for it in items if isinstance(it[0], O.Background)]
    packed = O.pack_elements(mags) if mags else None
    model = N.Model(packed) if mags else None
    cur = O.pack_currents(curs) if curs else None
    ent = (_gen[0], (mags, packed, model, cur, bck))
    if key in _cache and _cache[key][1][2] is not None:
        _cache[key][1][2].close()
    _cache[key] = ent
    return ent[1]


def _component(v, comp):
    comp = comp.lower()
    if len(comp) == 1:
        return list(map(float, v))
    idx = {"x": 0, "y": 1, "z": 2}
    out = [float(v[idx[c]]) for c in comp[1:]]
    return out[0] if len(out) == 1 else out


def Fld(obj, comp, P, use_gpu=True, precision="double"):
    """Field component(s) comp ('b','bx','h','hz',...) at one point or a list
    of points (radpy.cpp:2389-2456).  precision 'single', 'float' or 'fp32'
    with use_gpu evaluates the B integrals in FP32 (FP64 accumulation), as the
    reference's use_gpu code 2; 'double' and 'fp64' (and use_gpu=False, the
    reference's FP64 host path) evaluate them in FP64; H is always FP64.  Any
    other precision string raises the reference's argument error."""
    fp32 = False
    if precision is not None and precision != "":
        if not isinstance(precision, str):
            raise RuntimeError("Incorrect function arguments: Fld")
        if precision in ("single", "float", "fp32"):
            fp32 = bool(use_gpu)
        elif precision not in ("double", "fp64"):
            raise RuntimeError("Incorrect function arguments: Fld: precision must be 'double' "
                               "or 'single'")
    comp = str(comp).lower()
    if comp[0] not in "bh":
        raise NotImplementedError("only B and H components are supported")
    pts = np.asarray(P, dtype=np.float64)
    single = pts.ndim == 1
    pts = np.ascontiguousarray(pts.reshape(-1, 3))
    mags, packed, model, cur, bck = _field_model(obj)
    B = np.zeros_like(pts)
    H = np.zeros_like(pts)
    if model is not None:
        M = np.stack([o.M for (o, _k, _s) in mags])
        if comp[0] == "b" and fp32:
            # FP32 integrals for B only, as the reference's use_gpu code 2
            b, _h = model.field(M, pts, want_h=False, precision=N.RH_MIXED)
            B += b
        else:
            b, h = model.field(M, pts, precision=N.RH_FP64)
            B += b
            H += h
    if cur is not None:
        c = N.current_field(cur, pts)
        B += c
        H += c
    for o, _k, stack in bck:
        for t in O._expand(stack):
            v = t.s * (t.M @ o.B)
            B += v
            H += v
    src = B if comp[0] == "b" else H
    res = [_component(v, comp) for v in src]
    return res[0] if single else res


def FldLst(obj, comp, P1, P2, np_, arg="arg", start=0.0):
    """Field along the segment P1 -> P2 at np_ equally spaced points (np_ = 1
    with a finite P2 means 101), each preceded by its longitudinal position
    start + i*step when arg is 'arg' (radapl3.cpp:34-100); returned as the
    reference's Python layer shapes it (a flat list when each point has one
    value)."""
    a = str(arg).lower()
    if a not in ("arg", "noarg"):
        raise RuntimeError("Radia::Error034")
    p1 = np.asarray(_vec3(P1), dtype=np.float64)
    p2 = np.asarray(_vec3(P2), dtype=np.float64)
    n = int(np_)
    if n <= 0:
        raise RuntimeError("Incorrect argument: FldLst")
    if n == 1 and (p2 < 1e22).all():
        n = 101
    pts = [p1.copy()]
    argv = [float(start)]
    if n > 1:
        step = (1.0 / float(n - 1)) * (p2 - p1)
        ds = float(np.sqrt(step @ step))
        for _ in range(1, n):
            pts.append(pts[-1] + step)
            argv.append(argv[-1] + ds)
    vals = Fld(obj, comp, np.array(pts))
    rows = []
    for i, v in enumerate(vals):
        row = [argv[i]] if a == "arg" else []
        row += list(v) if isinstance(v, list) else [v]
        rows.append(row)
    if n == 1:
        flat = rows[0]
        return flat[0] if len(flat) == 1 else flat
    return [r[0] if len(r) == 1 else r for r in rows]


def _face_planes(o):
    """Outward face planes (n, d: n.x <= d inside) of a block or convex
    polyhedron in its own coordinates."""
    if isinstance(o, O.RecMag):
        c, h = np.asarray(o.center), 0.5 * np.asarray(o.dims)
        out = []
        for a in range(3):
            n = np.zeros(3)
            n[a] = 1.0
            out += [(n, c[a] + h[a]), (-n, -(c[a] - h[a]))]
        return out
    V = o.verts
    c = V.mean(0)
    out = []
    for f in o.faces:
        p0, p1, p2 = V[f[0]], V[f[1]], V[f[2]]
        n = np.cross(p1 - p0, p2 - p0)
        k = 2
        while np.linalg.norm(n) == 0.0 and k + 1 < len(f):
            k += 1
            n = np.cross(p1 - p0, V[f[k]] - p0)
        n = n / np.linalg.norm(n)
        if n @ (c - p0) > 0:
            n = -n
        out.append((n, float(n @ p0)))
    return out


def _line_breaks(obj, p0, u):
    """Parameters t where the line p0 + t u crosses an element surface (the
    field jumps only there)."""
    ts = []
    for o, _k, stack in O.walk(_reg, int(obj)):
        if not isinstance(o, (O.RecMag, O.Polyhedron)) or getattr(o, "is_current", False):
            continue
        planes = _face_planes(o)
        for t in O._expand(stack):
            # the copy's point map x -> M x + V: planes n.x <= d become
            # (M^-T n).x <= d + (M^-T n).V in the lab frame
            Mi = np.linalg.inv(t.M)
            for n, d in planes:
                nl = Mi.T @ n
                dl = d + nl @ t.V
                den = nl @ u
                if abs(den) < 1e-300:
                    continue
                tt = (dl - nl @ p0) / den
                x = p0 + tt * u
                xl = Mi @ (x - t.V)
                tol = 1e-9 * (1.0 + float(np.abs(xl).max()))
                if all(m @ xl <= e + tol for m, e in planes):
                    ts.append(float(tt))
    return sorted(set(ts))


_GL_X, _GL_W = np.polynomial.legendre.leggauss(20)


def _integrate(pieces, tol_rel=1e-12, max_level=9):
    """Sum of the integrals over the pieces (a, b, g, f): the vector function
    f at x = g(s) times the weight g'(s) over s in [a, b]; composite 20-point
    Gauss-Legendre, panels doubled per piece until two levels agree."""
    level = [0] * len(pieces)
    prev = [None] * len(pieces)
    done = [False] * len(pieces)
    total = None
    for _ in range(max_level + 1):
        todo = [i for i in range(len(pieces)) if not done[i]]
        if not todo:
            break
        vals = {}
        for fn in {id(pieces[i][3]): pieces[i][3] for i in todo}.values():
            vals.update(_integrate_level([i for i in todo if pieces[i][3] is fn], pieces, level, fn))
        cur = [vals[i] if i in vals else prev[i] for i in range(len(pieces))]
        total = np.sum(cur, axis=0)
        scale = np.abs(total).max() + 1e-300
        for i in todo:
            if prev[i] is not None and np.abs(vals[i] - prev[i]).max() <= tol_rel * scale:
                done[i] = True
            prev[i] = vals[i]
            level[i] += 1
    return total


def _integrate_level(todo, pieces, level, f):
    """Integrals of f over the pieces listed in todo, each split into
    2^level panels of 20 Gauss-Legendre nodes, with one call of f."""
    nodes, wts, owner = [], [], []
    for i in todo:
        a, b, g, _fn = pieces[i]
        m = 2 ** level[i]
        edges = np.linspace(a, b, m + 1)
        for k in range(m):
            lo, hi = edges[k], edges[k + 1]
            sN = 0.5 * (hi - lo) * _GL_X + 0.5 * (hi + lo)
            x, jac = g(sN)
            nodes.append(x)
            wts.append(0.5 * (hi - lo) * _GL_W * jac)
            owner.append(np.full(len(sN), i))
    X = np.concatenate(nodes)
    Wt = np.concatenate(wts)
    own = np.concatenate(owner)
    F = np.asarray(f(X))
    vals = {}
    for i in todo:
        sel = own == i
        vals[i] = (F[sel] * Wt[sel, None]).sum(0)
    return vals


def FldInt(obj, kind, comp, P1, P2):
    """Field integral along the straight line through P1 and P2 in T*mm:
    'fin' from P1 to P2, 'inf' over the whole line (radpy.cpp FldInt,
    radapl3.cpp:356-381).  comp: 'ib' or 'ih' with optional components
    ('ibx', 'ihyz', ...).  Computed from the device field: the line is split
    where it crosses element surfaces and at distances R 2^j (j = 0..16) from
    the centre of the sources' bounding sphere (radius R), each piece
    integrated by composite Gauss-Legendre with panel doubling to 1e-12 of the
    integral's scale.  The two tails of an infinite line beyond 65536 R
    (about 1e-10 of the integral) are mapped to a finite interval
    (t = t0 + L u/(1 - u)) and take the field of the total dipole moment,
    exact there to R/t (with current sources present, the device field)."""
    k = str(kind).lower()
    c = str(comp).lower()
    if k not in ("inf", "fin") or len(c) < 2 or c[0] != "i" or c[1] not in "bh":
        raise RuntimeError("Radia::Error033")
    p1 = np.asarray(_vec3(P1), dtype=np.float64)
    p2 = np.asarray(_vec3(P2), dtype=np.float64)
    seg = p2 - p1
    length = float(np.linalg.norm(seg))
    if not length > 0.0:
        raise RuntimeError("Incorrect argument: FldInt")
    u = seg / length
    which = c[1]
    idx = [{"x": 0, "y": 1, "z": 2}[ch] for ch in c[2:]] if len(c) > 2 else [0, 1, 2]

    def f(X):
        v = np.asarray(Fld(obj, which, X.reshape(-1, 3)), dtype=np.float64).reshape(-1, 3)
        return v

    brk = _line_breaks(obj, p1, u)
    # geometric breakpoints about the sources' bounding sphere, so that every
    # panel spans a shell over which the field varies smoothly
    pts_src = []
    for o, _k, stack in O.walk(_reg, int(obj)):
        if isinstance(o, O.RecMag):
            cc, hh = np.asarray(o.center), 0.5 * np.asarray(o.dims)
            V = np.array([cc + hh * np.array([sx, sy, sz]) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        elif isinstance(o, O.Polyhedron):
            V = o.verts
        else:
            continue
        for t in O._expand(stack):
            pts_src.append(V @ t.M.T + t.V)
    if pts_src:
        A = np.concatenate(pts_src)
        ctr = 0.5 * (A.min(0) + A.max(0))
        R = max(0.5 * float(np.linalg.norm(A.max(0) - A.min(0))), 1e-9 * (1.0 + length))
    else:
        ctr, R = 0.5 * (p1 + p2), 0.5 * length
    tc = float((ctr - p1) @ u)
    far = 65536.0
    geo = [tc + sgn * R * 2.0 ** j for j in range(0, 17) for sgn in (-1.0, 1.0)]
    pieces = []
    # beyond far * R the field is that of the total dipole moment to R/t
    # (the exact face formulas lose everything to cancellation far away)
    mom = np.zeros(3)
    has_cur = False
    for o, _k, stack in O.walk(_reg, int(obj)):
        if getattr(o, "is_current", False) or isinstance(o, (O.ArcCur,)):
            has_cur = True
        if isinstance(o, (O.RecMag, O.Polyhedron)) and not getattr(o, "is_current", False):
            if isinstance(o, O.RecMag):
                vol = float(np.prod(o.dims))
            else:
                V = o.verts
                cv = V.mean(0)
                vol = sum(abs(np.dot(V[f_[0]] - cv, np.cross(V[f_[j]] - cv, V[f_[j + 1]] - cv))) / 6.0
                          for f_ in o.faces for j in range(1, len(f_) - 1))
            for t in O._expand(stack):
                mom += vol * t.s * (t.M @ np.asarray(o.M))

    def f_dip(X):
        r = X - ctr[None, :]
        rn = np.linalg.norm(r, axis=1)
        e = r / rn[:, None]
        return (3.0 * (e @ mom)[:, None] * e - mom[None, :]) / (4.0 * np.pi * rn[:, None] ** 3)

    f_tail = f if has_cur else f_dip

    def lin(a, b):
        return (a, b, lambda s: (p1[None, :] + s[:, None] * u[None, :], np.ones_like(s)), f)

    if k == "fin":
        ts = sorted({0.0, length} | {t for t in brk + geo if 0.0 < t < length})
        for a, b in zip(ts[:-1], ts[1:]):
            if b > a:
                pieces.append(lin(a, b))
    else:
        t_lo = min([0.0, tc - R * far] + brk)
        t_hi = max([length, tc + R * far] + brk)
        ts = sorted({t_lo, t_hi} | {t for t in brk + geo if t_lo < t < t_hi})
        for a, b in zip(ts[:-1], ts[1:]):
            if b > a:
                pieces.append(lin(a, b))
        L = t_hi - t_lo
        pieces.append((0.0, 1.0, lambda s: (p1[None, :] + (t_hi + L * s / (1.0 - s))[:, None] * u[None, :],
                                            L / (1.0 - s) ** 2), f_tail))
        pieces.append((0.0, 1.0, lambda s: (p1[None, :] + (t_lo - L * s / (1.0 - s))[:, None] * u[None, :],
                                            L / (1.0 - s) ** 2), f_tail))
    tot = _integrate(pieces)
    out = [float(tot[i]) for i in idx]
    return out[0] if len(out) == 1 else out


def UtiMPI(*args):
    """MPI is not part of this build (one process per GPU, one GPU here):
    reports rank 0 of 1."""
    return [0, 1]


def ObjDrwOpenGL(*args, **kw):
    """3D viewer: graphics are not part of this build; does nothing."""
    return 0


# ------------------------------------------------------------------ utilities
def FldLenRndSw(onoff):
    """Length randomization switch.  The device formulas need no random
    jitter (edge/plane limits are evaluated in cancellation-free forms), so
    this only records the state."""
    s = str(onoff).lower()
    if s not in ("on", "off"):
        raise RuntimeError("Radia::Error043")
    _rand[0] = s
    return 1 if s == "on" else 0


def UtiDel(k):
    o = _reg.pop(int(k), None)
    if isinstance(o, _Interaction):
        o.matrix.close()
        o.model.close()
    _bump()
    return 0


def UtiDelAll():
    for k in list(_reg):
        UtiDel(k)
    for ent in _cache.values():
        if ent[1][2] is not None:
            ent[1][2].close()
    _cache.clear()
    _next[0] = 1
    _bump()
    return 0


def _backend_name(code):
    # radpy.cpp:3237-3276: 1 'gpu', 0 'cpu', otherwise 'none'; a matrix with
    # streamed rows (code 2) is serviced by the device too
    return "gpu" if code in (1, 2) else ("cpu" if code == 0 else "none")


def UtiFldLastBackend():
    """Backend of the most recent field evaluation (rh_last_backend(1))."""
    return _backend_name(N.last_backend(1))


def UtiAsmLastBackend():
    """Backend of the most recent interaction-matrix assembly (rh_last_backend(0))."""
    return _backend_name(N.last_backend(0))


def UtiGpuFallback(mode=None):
    """Sets and/or reports the library's fallback policy for a matrix beyond
    the device budget (rh_set_fallback; radpy.cpp:2919-2945): 'cpu'
    (default) and 'gpu_streaming' stream the rows that do not fit from host
    memory, 'break' raises Radia::Error602.  Names are matched exactly."""
    if mode is not None:
        m = mode if isinstance(mode, str) else None
        if m not in N.FALLBACK_NAMES:
            raise RuntimeError("UtiGpuFallback: expected 'cpu', 'gpu_streaming' or 'break'")
        N.set_fallback(N.FALLBACK_NAMES.index(m))
    return N.FALLBACK_NAMES[N.get_fallback()]


def UtiVer():
    return "radiacuda_amd 0.1 (gfx950)"
