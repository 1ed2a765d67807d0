"""Seeded synthetic models for the benchmark configurations, built through the
drop-in API (no files, no network).

c_yoke: C-shaped dipole yoke of tetrahedral ObjPolyhdr from a structured
hexahedral subdivision (5 tetrahedra per cell, alternating so that the mesh is
conforming), interior vertices jittered by +-jitter of the cell size, steel
MatSatIsoFrm([2118,1.362],[63.06,0.2605],[17.138,0.4917]) and two racetrack
coils around the back leg; field points on a regular grid in the gap.
"""
import numpy as np

from . import rad

STEEL = ([2118.0, 1.362], [63.06, 0.2605], [17.138, 0.4917])
TET_FACES = [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]]
# 5-tetrahedron split of a cell (corner index 4i+2j+k) and its mirror image
FIVE_EVEN = [[1, 2, 4, 7], [0, 1, 2, 4], [3, 1, 2, 7], [5, 1, 4, 7], [6, 2, 4, 7]]
FIVE_ODD = [[0, 3, 5, 6], [1, 0, 3, 5], [2, 0, 3, 6], [4, 0, 5, 6], [7, 3, 5, 6]]


def c_section(scale=1):
    """Cells (ix, iz) of the C cross-section: back leg 4x20, two arms 10x4."""
    cells = set()
    for ix in range(4 * scale):
        for iz in range(20 * scale):
            cells.add((ix, iz))
    for ix in range(4 * scale, 14 * scale):
        for iz in list(range(4 * scale)) + list(range(16 * scale, 20 * scale)):
            cells.add((ix, iz))
    return sorted(cells)


def c_yoke(ny=10, h=10.0, jitter=0.1, seed=0, j_coil=5.0, scale=1):
    """Returns (top, iron, coils, info).  ny * 160 * scale^2 cells, 5 tets each:
    ny=10, scale=1 gives 8,000 tetrahedra (4.0 mm... cell size h in mm)."""
    rng = np.random.default_rng(seed)
    hh = h / scale
    sec = c_section(scale)
    cells = [(ix, iy, iz) for (ix, iz) in sec for iy in range(ny)]
    cellset = set(cells)
    # vertex lattice with jitter on interior vertices (all 8 adjacent cells present)
    verts = {}

    def vert(i, j, k):
        key = (i, j, k)
        v = verts.get(key)
        if v is None:
            v = np.array([i * hh, j * hh, k * hh], dtype=np.float64)
            inner = all((i - a, j - b, k - c) in cellset for a in (0, 1) for b in (0, 1)
                        for c in (0, 1))
            if inner:
                v = v + rng.uniform(-jitter * hh, jitter * hh, 3)
            verts[key] = v
        return v

    ids = []
    for (i, j, k) in cells:
        c = np.array([vert(i + a, j + b, k + d) for a in (0, 1) for b in (0, 1) for d in (0, 1)])
        split = FIVE_EVEN if (i + j + k) % 2 == 0 else FIVE_ODD
        for t in split:
            ids.append(rad.ObjPolyhdr(c[t].tolist(), TET_FACES, [0, 0, 0]))
    iron = rad.ObjCnt(ids)
    rad.MatApl(iron, rad.MatSatIsoFrm(*STEEL))
    L = 4 * h          # back-leg thickness along x
    W = ny * h         # depth along y
    coils = []
    for zc in (0.3 * 20 * h, 0.7 * 20 * h):
        coils.append(rad.ObjRaceTrk([0.5 * L, 0.5 * W, zc], [0.5 * h, 2.5 * h], [L, W], 3 * h, 10,
                                    j_coil, "man", "z"))
    coil = rad.ObjCnt(coils)
    top = rad.ObjCnt([iron, coil])
    info = dict(n_tets=len(ids), cell=h, gap=(4 * h, 14 * h, 0.0, W, 4 * h, 16 * h))
    return top, iron, coil, info


def gap_grid(info, n=100, margin=0.1):
    """n^3 regular grid inside the gap, inset by margin of each extent."""
    x0, x1, y0, y1, z0, z1 = info["gap"]
    ax = []
    for a, b in ((x0, x1), (y0, y1), (z0, z1)):
        d = (b - a) * margin
        ax.append(np.linspace(a + d, b - d, n))
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


HEX_FACES = [[1, 2, 4, 3], [5, 6, 8, 7], [1, 2, 6, 5], [3, 4, 8, 7], [1, 3, 7, 5], [2, 4, 8, 6]]


def undulator(n_blocks=2000, period=20.0, gap=10.0, width=40.0, height=20.0, br=1.2,
              ksi=(0.06, 0.17), seed=0, n_points=4096):
    """Planar pure-permanent-magnet undulator: two jaws of hexahedral
    ObjPolyhdr blocks (4 per period, magnetization turning 90 deg per block,
    the lower jaw mirrored), MatLin(ksi, Br), field points on the axis over
    the central half of the device, and a seeded target By (the ideal
    sinusoid with 1 % noise).  Returns (top, info)."""
    rng = np.random.default_rng(seed)
    per_jaw = n_blocks // 2
    L = per_jaw * period / 4.0
    bz = period / 4.0
    z0 = -0.5 * L
    dirs = [(0, 1, 0), (0, 0, -1), (0, -1, 0), (0, 0, 1)]
    ids = []
    for jaw, sy in ((0, 1.0), (1, -1.0)):
        for k in range(per_jaw):
            zc = z0 + (k + 0.5) * bz
            yc = sy * (0.5 * gap + 0.5 * height)
            d = np.array(dirs[k % 4], dtype=float)
            if jaw == 1:
                d = np.array([d[0], d[1], -d[2]])
            c = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
            c = c * np.array([width, height, bz]) + np.array([0.0, yc, zc])
            ids.append(rad.ObjPolyhdr(c.tolist(), HEX_FACES, list(br * d)))
    top = rad.ObjCnt(ids)
    rad.MatApl(top, rad.MatLin(list(ksi), br))
    z = np.linspace(-0.25 * L, 0.25 * L, n_points)
    pts = np.stack([np.zeros_like(z), np.zeros_like(z), z], 1)
    b0 = 2.0 * br * np.exp(-np.pi * gap / period) * (1 - np.exp(-2 * np.pi * height / period))
    target = b0 * np.cos(2 * np.pi * z / period) * (1 + 0.01 * rng.normal(size=n_points))
    return top, dict(points=pts, target=target, period=period, length=L)


def iron_cube(n_points=10000, j=3.0):
    """40 mm iron cube (ObjRecMag subdivided 8x8x8, MatSatIsoFrm steel) inside
    a square loop of four ObjRecCur bars (10x10 mm section, 60 mm loop, J
    A/mm^2); field points on a tilted line through the centre, offset so that
    no point lies on a block plane.  Returns (top, iron, coil, points)."""
    cube = rad.ObjRecMag([0, 0, 0], [40, 40, 40], [0, 0, 0])
    rad.ObjDivMag(cube, [8, 8, 8])
    rad.MatApl(cube, rad.MatSatIsoFrm(*STEEL))
    bars = [rad.ObjRecCur([0, -30, 0], [70, 10, 10], [j, 0, 0]),
            rad.ObjRecCur([30, 0, 0], [10, 50, 10], [0, j, 0]),
            rad.ObjRecCur([0, 30, 0], [70, 10, 10], [-j, 0, 0]),
            rad.ObjRecCur([-30, 0, 0], [10, 50, 10], [0, -j, 0])]
    coil = rad.ObjCnt(bars)
    top = rad.ObjCnt([cube, coil])
    t = np.linspace(-1.0, 1.0, n_points)
    d = np.array([50.0, 41.0, 33.0])
    off = np.array([0.123, -0.271, 0.377])
    pts = off + t[:, None] * d
    return top, cube, coil, pts


def halbach(n_seg=16, n_az=10, n_rad=5, n_ax=25, r_in=20.0, r_out=40.0, length=50.0, br=1.2,
            ksi=(0.06, 0.17), err_deg=1.0, seed=0, n_points=2000000):
    """Halbach ring of hexahedral ObjPolyhdr blocks (annular sectors with chord
    faces): n_seg segments around the ring, each n_az x n_rad x n_ax blocks,
    magnetization in the ring plane turning 90 deg per segment with seeded
    +-err_deg errors per block, MatLin(ksi, Br); field points uniform in the
    bore (r < 0.75 r_in, |z| < 0.4 length).  Returns (top, points)."""
    rng = np.random.default_rng(seed)
    ids = []
    nphi = n_seg * n_az
    rs = np.linspace(r_in, r_out, n_rad + 1)
    zs = np.linspace(-0.5 * length, 0.5 * length, n_ax + 1)
    for ip in range(nphi):
        p0, p1 = 2 * np.pi * ip / nphi, 2 * np.pi * (ip + 1) / nphi
        seg = ip // n_az
        for ir in range(n_rad):
            for iz in range(n_ax):
                ang = np.deg2rad(90.0 * seg + rng.uniform(-err_deg, err_deg))
                m = br * np.array([np.cos(ang), np.sin(ang), 0.0])
                c = []
                for r in (rs[ir], rs[ir + 1]):
                    for p in (p0, p1):
                        for z in (zs[iz], zs[iz + 1]):
                            c.append([r * np.cos(p), r * np.sin(p), z])
                ids.append(rad.ObjPolyhdr(c, HEX_FACES, list(m)))
    top = rad.ObjCnt(ids)
    rad.MatApl(top, rad.MatLin(list(ksi), br))
    u = rng.uniform(size=(n_points, 3))
    r = 0.75 * r_in * np.sqrt(u[:, 0])
    p = 2 * np.pi * u[:, 1]
    z = 0.4 * length * (2 * u[:, 2] - 1)
    pts = np.stack([r * np.cos(p), r * np.sin(p), z], 1)
    return top, pts
