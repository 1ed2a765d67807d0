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
