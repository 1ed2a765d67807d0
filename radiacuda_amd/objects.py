"""Host object model behind the ``rad`` drop-in API.

Field sources (blocks, polyhedra, arc and block currents), containers, space
transforms with multiplicities, and materials; plus the flattening that turns
an object tree into the flat arrays the C ABI takes.

The flattening follows Radia's own conventions:
  * a transform list is kept per object with the most recently applied one
    first (radg3d.h:871-875); a container's transforms are outside its
    members' (radintrc.h:295-300, radintrc.cpp:526-553);
  * a transform of multiplicity m > 1 contributes its powers 0..m-1, one of
    multiplicity 1 is always applied (radg3d.cpp:57-115);
  * copy 0 of an element (all symmetry powers zero) is its main transform,
    which defines the collocation point and the row transform of the
    interaction matrix (radintrc.cpp:557-573, 852);
  * elements with a material are relaxed, every other source is external
    (radintrc.cpp:248-270).
"""
import itertools
import math

import numpy as np

# ---------------------------------------------------------------- transforms


class Trf:
    """Affine map x -> M x + V acting on field vectors as h -> s M h (gmtrans.h:30-85)."""

    __slots__ = ("M", "V", "s")

    def __init__(self, M=None, V=None, s=1.0):
        self.M = np.eye(3) if M is None else np.asarray(M, dtype=np.float64).reshape(3, 3)
        self.V = np.zeros(3) if V is None else np.asarray(V, dtype=np.float64).reshape(3)
        self.s = float(s)

    def compose(self, inner):
        """self o inner (TrProduct in gmtrans.h)."""
        return Trf(self.M @ inner.M, self.M @ inner.V + self.V, self.s * inner.s)

    def power(self, k):
        t = Trf()
        for _ in range(k):
            t = t.compose(self)
        return t

    @staticmethod
    def translation(v):
        return Trf(V=v)

    @staticmethod
    def rotation(p, axis, phi):
        """Right-handed rotation by phi about the axis through p (gmtrans.cpp:16-46)."""
        a = np.asarray(axis, dtype=np.float64)
        a = a / np.linalg.norm(a)
        c, s = math.cos(phi), math.sin(phi)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R = c * np.eye(3) + s * K + (1 - c) * np.outer(a, a)
        p = np.asarray(p, dtype=np.float64)
        return Trf(R, (np.eye(3) - R) @ p)

    @staticmethod
    def plane_symmetry(p, n):
        n = np.asarray(n, dtype=np.float64)
        n = n / np.linalg.norm(n)
        M = np.eye(3) - 2.0 * np.outer(n, n)
        return Trf(M, 2.0 * np.dot(np.asarray(p, dtype=np.float64), n) * n)

    @staticmethod
    def field_inversion():
        return Trf(s=-1.0)


# ---------------------------------------------------------------- objects


class Obj:
    kind = "obj"

    def __init__(self):
        self.trfs = []  # [(Trf, mult)], most recently applied first

    def add_trf(self, trf, mult):
        self.trfs.insert(0, (trf, int(mult)))


class RecMag(Obj):
    """Rectangular block with uniform magnetization or current density."""

    kind = "recmag"

    def __init__(self, center, dims, M=(0, 0, 0), J=(0, 0, 0)):
        super().__init__()
        self.center = np.asarray(center, dtype=np.float64).reshape(3)
        self.dims = np.asarray(dims, dtype=np.float64).reshape(3)
        self.M = np.asarray(M, dtype=np.float64).reshape(3).copy()
        self.J = np.asarray(J, dtype=np.float64).reshape(3).copy()
        self.mat = None

    @property
    def is_current(self):
        return bool(np.any(self.J != 0))


class Polyhedron(Obj):
    kind = "poly"

    def __init__(self, verts, faces, M=(0, 0, 0)):
        super().__init__()
        self.verts = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
        self.faces = [list(map(int, f)) for f in faces]  # 0-based
        self.M = np.asarray(M, dtype=np.float64).reshape(3).copy()
        self.mat = None
        if len(self.faces) < 4 or len(self.verts) < 4:
            raise ValueError("a polyhedron needs at least 4 vertices and 4 faces")
        for f in self.faces:
            if len(f) < 3 or min(f) < 0 or max(f) >= len(self.verts):
                raise ValueError("face vertex index out of range")


def cut_convex(verts, faces, normal, offset, tol):
    """Splits a convex polyhedron (vertices, face loops) by the plane
    normal . x = offset.  Returns (below, above): each a (vertices, faces)
    pair, or None when that side holds no volume.  Vertices within tol of the
    plane count as on it; the cut adds one vertex per crossed edge and a cap
    face on each side (the outward orientation of every face is settled by
    the packer).  The pieces of radTPolyhedron::CutItself
    (cpp/src/core/radvlpgn.cpp:1330), computed here by clipping each face
    loop against the half-spaces."""
    V = np.asarray(verts, dtype=np.float64)
    nrm = np.asarray(normal, dtype=np.float64)
    sd = V @ nrm - offset
    sd[np.abs(sd) <= tol] = 0.0
    if (sd >= 0.0).all():
        return None, (V, [list(f) for f in faces])
    if (sd <= 0.0).all():
        return (V, [list(f) for f in faces]), None
    out = list(V)
    side = list(np.sign(sd))
    made = {}

    def crossing(i, j):
        key = (min(i, j), max(i, j))
        if key not in made:
            t = sd[i] / (sd[i] - sd[j])
            out.append(V[i] + t * (V[j] - V[i]))
            side.append(0.0)
            made[key] = len(out) - 1
        return made[key]

    lower, upper, cap = [], [], set()
    for f in faces:
        lo, hi = [], []
        m = len(f)
        for k in range(m):
            i, j = f[k], f[(k + 1) % m]
            if sd[i] <= 0.0:
                lo.append(i)
            if sd[i] >= 0.0:
                hi.append(i)
            if sd[i] == 0.0:
                cap.add(i)
            if sd[i] * sd[j] < 0.0:
                x = crossing(i, j)
                lo.append(x)
                hi.append(x)
                cap.add(x)
        if len(lo) >= 3 and any(side[v] < 0 for v in lo):
            lower.append(lo)
        if len(hi) >= 3 and any(side[v] > 0 for v in hi):
            upper.append(hi)
    cap = sorted(cap)
    if len(cap) >= 3:
        P = np.array([out[i] for i in cap])
        c = P.mean(0)
        u = np.cross(nrm, [1.0, 0.0, 0.0] if abs(nrm[0]) < 0.9 else [0.0, 1.0, 0.0])
        u /= np.linalg.norm(u)
        w = np.cross(nrm, u)
        ang = np.arctan2((P - c) @ w, (P - c) @ u)
        loop = [cap[i] for i in np.argsort(ang)]
        lower.append(loop)
        upper.append(loop[::-1])

    def compact(fs):
        used = sorted({v for f in fs for v in f})
        remap = {v: k for k, v in enumerate(used)}
        return np.array([out[v] for v in used]), [[remap[v] for v in f] for f in fs]

    return compact(lower), compact(upper)


class LayerError(ValueError):
    """A layer structure the reference refuses: code 111 (a non-convex layer
    polygon) or 63 (any other improper layer definition)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _ccw_layer(pts, tol):
    """Layer points as a counter-clockwise convex loop (1 or 2 points are
    returned as given).  LayerError 111 for a non-convex polygon, 63 for a
    degenerate one."""
    P = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
    if len(P) < 3:
        return P
    x, y = P[:, 0], P[:, 1]
    area2 = float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))
    size = float(np.ptp(P, axis=0).max())
    if abs(area2) <= tol * size * size:
        raise LayerError(63, "degenerate layer polygon")
    if area2 < 0:
        P = P[::-1].copy()
    e = np.roll(P, -1, axis=0) - P
    cr = e[:, 0] * np.roll(e, -1, axis=0)[:, 1] - e[:, 1] * np.roll(e, -1, axis=0)[:, 0]
    if (cr < -tol * size * size).any():
        raise LayerError(111, "non-convex layer polygon")
    return P


def layer_segment(pa, za, pb, zb, tol=1e-9):
    """Convex polyhedron spanned by two parallel convex layers: polygon pa at
    z = za and polygon pb at z = zb, each of 1 or 2 points (an apex or a
    ridge) or a convex polygon.  Returns (verts, faces): the points of pa then
    those of pb, the layer polygons as faces when they have 3 or more points,
    and the lateral faces of the hull: each layer edge with the other layer's
    supporting vertex in the edge's outward direction (a triangle), or with
    the other layer's parallel edge (a quadrilateral).  Radia's segment
    construction (radTApplication::SetUpOnePolyhedronSegment,
    cpp/src/core/radapl4.cpp:519-700) builds the same faces by searching
    each edge's lowest point on the other layer."""
    A = _ccw_layer(pa, tol)
    B = _ccw_layer(pb, tol)
    m, k = len(A), len(B)
    if m < 3 and k < 3 and not (m == 2 and k == 2):
        raise LayerError(63, "two consecutive layers of fewer than 3 points must both hold 2 points")
    size = max(float(np.ptp(np.concatenate([A, B]), axis=0).max()), abs(zb - za))
    if m == 2 and k == 2:
        ea, eb = A[1] - A[0], B[1] - B[0]
        if abs(ea[0] * eb[1] - ea[1] * eb[0]) <= tol * size * size:
            raise LayerError(63, "parallel ridges in consecutive layers")

    def edges(P, off):
        # (outward normal angle, start index, end index) of each edge; a
        # two-point layer has both directions of its one segment
        out = []
        n = len(P)
        if n < 2:
            return out
        for i in range(n if n > 2 else 2):
            j = (i + 1) % n
            e = P[j] - P[i]
            out.append((math.atan2(-e[0], e[1]), off + i, off + j, i))
        return out

    def support(P, ang):
        u = np.array([math.cos(ang), math.sin(ang)])
        return int(np.argmax(P @ u))

    verts = np.array([[x, y, za] for x, y in A] + [[x, y, zb] for x, y in B])
    faces = []
    if m >= 3:
        faces.append(list(range(m)))
    if k >= 3:
        faces.append(list(range(m, m + k)))
    ea, eb = edges(A, 0), edges(B, m)
    used_b = set()
    angtol = 1e-9
    for (ang, i0, i1, _i) in ea:
        par = None
        for (bang, j0, j1, j) in eb:
            d = math.remainder(bang - ang, 2.0 * math.pi)
            if abs(d) <= angtol and j not in used_b:
                par = (j0, j1, j)
                break
        if par is not None:
            used_b.add(par[2])
            faces.append([i0, i1, par[1], par[0]])
        else:
            faces.append([i0, i1, m + support(B, ang)])
    for (bang, j0, j1, j) in eb:
        if j in used_b:
            continue
        faces.append([j0, j1, support(A, bang)])
    return verts, faces


def division_planes(lo, hi, count, ratio):
    """Offsets of the count-1 cutting planes between the extreme vertex
    projections lo < hi, the piece widths in geometric progression with
    last/first = ratio (radTPolyhedron::DeterminePointsOnCuttingPlanes,
    cpp/src/core/radvlpgn.cpp:2386-2420)."""
    if count <= 1:
        return []
    q0 = ratio ** (1.0 / (count - 1))
    den = ratio * q0 - 1.0
    a1 = (q0 - 1.0) / den if abs(den) > 1e-13 else 1.0 / count
    out, tau, dt = [], a1, a1
    for _ in range(count - 1):
        out.append(lo + tau * (hi - lo))
        dt *= q0
        tau += dt
    return out


class ArcCur(Obj):
    """Arc of rectangular section carrying azimuthal current density (radarccu.h)."""

    kind = "arccur"

    def __init__(self, center, rmin, rmax, phimin, phimax, height, nseg, j):
        super().__init__()
        self.center = np.asarray(center, dtype=np.float64).reshape(3)
        self.rmin, self.rmax = float(rmin), float(rmax)
        self.phimin, self.phimax = float(phimin), float(phimax)
        self.height, self.nseg, self.j = float(height), int(nseg), float(j)


class Background(Obj):
    kind = "bckg"

    def __init__(self, B):
        super().__init__()
        self.B = np.asarray(B, dtype=np.float64).reshape(3)


class Container(Obj):
    kind = "cnt"

    def __init__(self, members):
        super().__init__()
        self.members = list(members)


# ---------------------------------------------------------------- materials

MAT_LIN_ISO, MAT_LIN_ANISO, MAT_NONLIN_FRM, MAT_NONLIN_TAB = 0, 1, 2, 3


class Material:
    """A material law (radmater.h).  ``finish`` binds an anisotropic linear
    material's easy axis to an element's magnetization (radmater.h:235-250)."""

    def __init__(self, kind, **kw):
        self.kind = kind
        self.p = kw

    def record(self, elem_M):
        """Per-element material record (rh_material layout, 16 doubles)."""
        r = np.zeros(16)
        r[0] = self.kind
        if self.kind == MAT_LIN_ISO:
            r[1] = self.p["ksi"]
            r[13:16] = self.p.get("mr", (0, 0, 0))
        elif self.kind == MAT_LIN_ANISO:
            kpar, kper = self.p["kpar"], self.p["kper"]
            mr = self.p.get("mr_vec")
            if mr is None:
                m = np.asarray(elem_M, dtype=np.float64)
                am = float(np.linalg.norm(m))
                if am < 1e-10:
                    raise RuntimeError("Radia::Error107: zero magnetization, cannot define "
                                       "the easy axis of a linear anisotropic material")
                mr = (self.p["mr"] / am) * m
            mr = np.asarray(mr, dtype=np.float64)
            amr = float(np.linalg.norm(mr))
            L = mr / amr if amr > 0 else np.array([0.0, 0.0, 1.0])
            K = kper * np.eye(3) + (kpar - kper) * np.outer(L, L)
            r[4:13] = K.ravel()
            r[13:16] = mr
        elif self.kind == MAT_NONLIN_FRM:
            ks, ms = self.p["ks"], self.p["ms"]
            r[1] = len(ks)
            r[2:2 + len(ms)] = ms
            r[5:5 + len(ks)] = ks
        elif self.kind == MAT_NONLIN_TAB:
            # [1] first node in the solve's table (set by pack_materials), [2] nodes
            r[1] = -1.0
            r[2] = len(self.p["x"])
        else:
            raise ValueError("unknown material kind %r" % self.kind)
        return r

    def table(self):
        """Flat (|H|, |M|, d|M|/d|H|) node triples of a tabulated law."""
        return np.stack([self.p["x"], self.p["y"], self.p["d"]], 1).ravel()


def _node_derivative(x, y, k):
    """Derivative at node k of the polynomial through the points (x, y), set to
    0 when it is negative on increasing data (radmater.cpp:341-455)."""
    n = len(x)
    der = 0.0
    for j in range(n):
        if j == k:
            der += y[j] * sum(1.0 / (x[k] - x[m]) for m in range(n) if m != k)
        else:
            w = 1.0 / (x[j] - x[k])
            for m in range(n):
                if m != j and m != k:
                    w *= (x[k] - x[m]) / (x[j] - x[m])
            der += y[j] * w
    inc = all(x[i + 1] > x[i] for i in range(n - 1)) and all(y[i + 1] >= y[i] for i in range(n - 1))
    return 0.0 if (inc and der < 0.0) else der


def tab_nodes(points):
    """Node table of a tabulated isotropic law M(H) from (H, M) points, as the
    reference builds it (radmater.h:436-451, radmater.cpp:460-609): (0, 0) is
    prepended unless the first point is at the origin; node derivatives come
    from 3- and 5-point interpolants in the reference's pattern (including its
    index use for 4 to 6 nodes) and are corrected so that each cubic Hermite
    segment over monotone data stays monotone.  Returns x, y, d."""
    pts = [(float(h), float(m)) for h, m in points]
    if pts[0][0] > 1e-10 and pts[0][1] > 1e-10:
        pts = [(0.0, 0.0)] + pts
    x = [p[0] for p in pts]
    y = [p[1] for p in pts]
    L = len(x)
    if L < 4:
        raise ValueError("a tabulated law needs at least 4 nodes (3 points above the origin)")
    d = [0.0] * L
    D3 = lambda t, k: _node_derivative(x[t:t + 3], y[t:t + 3], k)
    D5 = lambda t, k: _node_derivative(x[t:t + 5], y[t:t + 5], k)
    slope = lambda t: (y[t + 1] - y[t]) / (x[t + 1] - x[t])
    d[0], d[1], d[2] = D3(0, 0), D3(0, 1), D3(1, 1)
    w, t = 3, 1
    for _i in range(3, L - 4):
        d[w] = D5(t, 2)
        w += 1
        t += 1
    if L >= 7:
        d[w] = D5(t, 2)
        t += 2
        d[w + 1] = D3(t, 1)
        t += 1
        d[w + 2] = D3(t, 1)
        t += 1
        d[w + 3] = slope(t)
    elif L >= 6:
        d[w] = D5(t, 2)
        t += 2
        d[w + 1] = D3(t, 1)
        t += 1
        d[w + 2] = slope(t)
    elif L >= 5:
        d[w] = D3(t, 1)
        t += 1
        d[w + 1] = slope(t)
    else:
        d[w] = slope(t)
    # monotone correction, last segment first
    for i in range(L - 2, -1, -1):
        h = x[i + 1] - x[i]
        fm = y[i] - y[i + 1]
        p1, p2 = d[i], d[i + 1]
        c1 = p1
        c2 = -(3 * fm + (2 * p1 + p2) * h) / (h * h)
        c3 = (2 * fm + (p1 + p2) * h) / (h * h * h)
        disc = c2 * c2 - 3.0 * c1 * c3
        roots = []
        if c3 != 0.0 and disc >= 0.0:
            r = np.sqrt(disc)
            roots = [(-c2 - r) / (3.0 * c3), (-c2 + r) / (3.0 * c3)]
        mono = (y[i] < y[i + 1] and p1 > 0 and p2 > 0) or (y[i] > y[i + 1] and p1 < 0 and p2 < 0)
        inside = any(0.0 < z < h for z in roots)
        if mono and disc > 0.0 and inside:
            q = -3 * p2 * h ** 3 * (4 * fm + p2 * h)
            if q >= 0.0:
                sq = np.sqrt(q)
                base = -h * (6 * fm + p2 * h)
                fa = (base - sq) / (2 * h * h)
                fb = (base + sq) / (2 * h * h)
                d[i] = fa if abs(p1 - fa) < abs(p1 - fb) else fb
    return np.array(x), np.array(y), np.array(d)


def tab_eval(x, y, d, h):
    """|M| at |H| = h of a node table: cubic Hermite between nodes, y[0] below
    the first node, the last node's slope beyond the last (radmater.cpp:619-646)."""
    L = len(x)
    i = int(np.searchsorted(x, h, side="right")) - 1
    if i < 0:
        return float(y[0])
    if i >= L - 1:
        return float(y[-1] + (h - x[-1]) * d[-1])
    s = x[i + 1] - x[i]
    u = h - x[i]
    sl = (y[i + 1] - y[i]) / s
    c2 = (3 * sl - 2 * d[i] - d[i + 1]) / s
    c3 = (-2 * sl + d[i] + d[i + 1]) / (s * s)
    return float(y[i] + u * (d[i] + u * (c2 + u * c3)))


def pack_materials(items):
    """Material records [n, 16] and the node table their tabulated laws index
    (flat (x, y, d) triples; record [1] = first node)."""
    recs = []
    tables = []
    offsets = {}
    n_nodes = 0
    for (o, _k, _s) in items:
        r = o.mat.record(o.M)
        if o.mat.kind == MAT_NONLIN_TAB:
            key = id(o.mat)
            if key not in offsets:
                offsets[key] = n_nodes
                t = o.mat.table()
                tables.append(t)
                n_nodes += len(t) // 3
            r[1] = offsets[key]
        recs.append(r)
    tab = np.concatenate(tables) if tables else np.zeros(0)
    return np.stack(recs) if recs else np.zeros((0, 16)), tab


# ---------------------------------------------------------------- flattening


def _expand(trf_lists):
    """All copies of an element given its transform lists outer->inner.

    Returns a list of Trf; copy 0 has every symmetry power at zero."""
    factors = []
    for lst in trf_lists:
        for trf, m in lst:
            if m == 1:
                factors.append([trf])
            else:
                factors.append([trf.power(k) for k in range(m)])
    copies = []
    for combo in itertools.product(*factors) if factors else [()]:
        t = Trf()
        for f in combo:
            t = t.compose(f)
        copies.append(t)
    return copies


def walk(registry, key, trf_stack=()):
    """Yield (leaf_obj, leaf_key, transform lists outer->inner) for every leaf."""
    o = registry[key]
    stack = trf_stack + (o.trfs,) if o.trfs else trf_stack
    if isinstance(o, Container):
        for k in o.members:
            yield from walk(registry, k, stack)
    else:
        yield o, key, stack


class Flat:
    """Flat arrays of a model: relaxable elements and external sources."""


def flatten(registry, key, relax_only=False):
    """Split the tree under ``key`` into relaxable elements (with a material)
    and external sources, and build the flat element arrays."""
    rel, ext = [], []
    for o, k, stack in walk(registry, key):
        if isinstance(o, (RecMag, Polyhedron)) and not (isinstance(o, RecMag) and o.is_current):
            if o.mat is not None or not relax_only:
                (rel if o.mat is not None else ext).append((o, k, stack))
            else:
                ext.append((o, k, stack))
        else:
            ext.append((o, k, stack))
    return rel, ext


def pack_elements(items):
    """Flat element arrays (rh_model_desc layout) for magnetizable elements."""
    n = len(items)
    typ = np.zeros(n, np.int32)
    rec_c = np.zeros((n, 3))
    rec_d = np.zeros((n, 3))
    vert_off = np.zeros(n + 1, np.int64)
    face_off = np.zeros(n + 1, np.int64)
    copy_off = np.zeros(n + 1, np.int64)
    verts, face_len, face_idx = [], [], []
    cM, cV, cS = [], [], []
    M = np.zeros((n, 3))
    for e, (o, _k, stack) in enumerate(items):
        M[e] = o.M
        if isinstance(o, RecMag):
            typ[e] = 0
            rec_c[e] = o.center
            rec_d[e] = o.dims
            vert_off[e + 1] = vert_off[e]
            face_off[e + 1] = face_off[e]
        else:
            typ[e] = 1
            verts.append(o.verts)
            vert_off[e + 1] = vert_off[e] + len(o.verts)
            face_off[e + 1] = face_off[e] + len(o.faces)
            for f in o.faces:
                face_len.append(len(f))
                face_idx.extend(f)
        cps = _expand(stack)
        copy_off[e + 1] = copy_off[e] + len(cps)
        for t in cps:
            cM.append(t.M)
            cV.append(t.V)
            cS.append(t.s)
    return dict(
        type=typ, rec_c=rec_c, rec_d=rec_d, vert_off=vert_off,
        verts=np.concatenate(verts) if verts else np.zeros((0, 3)),
        face_off=face_off, face_len=np.asarray(face_len, np.int32),
        face_idx=np.asarray(face_idx, np.int32), copy_off=copy_off,
        cM=np.asarray(cM, dtype=np.float64).reshape(-1, 3, 3),
        cV=np.asarray(cV, dtype=np.float64).reshape(-1, 3),
        cS=np.asarray(cS, dtype=np.float64), M=M)


def pack_currents(items):
    """Flat arrays for current sources, each expanded over its copies.

    Kind 0: block current (centre, dims, J); kind 1: arc current.
    Every record is in its own frame with one copy transform."""
    recs = []
    for o, _k, stack in items:
        for t in _expand(stack):
            if isinstance(o, RecMag) and o.is_current:
                p = np.concatenate([o.center, o.dims, o.J, np.zeros(3)])
                recs.append((0, p, t))
            elif isinstance(o, ArcCur):
                p = np.array([o.center[0], o.center[1], o.center[2], o.rmin, o.rmax, o.phimin,
                              o.phimax, o.height, o.j, float(o.nseg), 0, 0])
                recs.append((1, p, t))
    kind = np.array([r[0] for r in recs], np.int32)
    par = np.array([r[1] for r in recs], dtype=np.float64).reshape(-1, 12)
    tM = np.array([r[2].M for r in recs], dtype=np.float64).reshape(-1, 3, 3)
    tV = np.array([r[2].V for r in recs], dtype=np.float64).reshape(-1, 3)
    tS = np.array([r[2].s for r in recs], dtype=np.float64)
    return dict(kind=kind, par=par, cM=tM, cV=tV, cS=tS)
