"""ObjMltExtPgn on the CPU: the layer-pair polyhedra are the convex hulls of
consecutive layers (volume and surface area against scipy's Qhull, every face
a supporting plane), layers of 1 and 2 points and of different vertex counts
are accepted, and improper layer structures raise the reference's errors
(radapl4.cpp:435-510, 1036-1064; radiobuf.cpp Error063, Error111).  The
reference's own example (examples/RADIA_Example04.py, a 15 x 15 polyhedral
sphere) is built and its interior field checked with the oracle."""
import math

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from oracle import oracle as OR
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from tests import models


def example04_sphere(r=1.0, n_phi=15, nz=15, M=(1, 0, 0)):
    """The slice structure of examples/RADIA_Example04.py:17-33."""
    dz = 2.0 * r / nz
    z = -r + dz
    dphi = 2.0 * math.pi / n_phi
    slices = [[[[0.0, 0.0]], -r]]
    for _i in range(1, nz):
        c = math.cos(math.asin(z / r))
        pg = [[r * c, 0.0]]
        phi = dphi
        for _k in range(1, n_phi):
            pg.append([r * math.cos(phi) * c, r * math.sin(phi) * c])
            phi += dphi
        slices.append([pg, z])
        z += dz
    slices.append([[[0.0, 0.0]], r])
    return rad.ObjMltExtPgn(slices, list(M))


def _check_hull(o):
    V = np.asarray(o.verts)
    hull = ConvexHull(V)
    c = V.mean(axis=0)
    vol = area = 0.0
    size = np.ptp(V, axis=0).max()
    for f in o.faces:
        P = V[f]
        nv = np.zeros(3)
        for i in range(len(f)):
            nv += np.cross(P[i], P[(i + 1) % len(f)])
        a = 0.5 * np.linalg.norm(nv)
        nv /= np.linalg.norm(nv)
        if np.dot(nv, P[0] - c) < 0:
            nv = -nv
        # a supporting plane: every vertex on the inner side
        assert (V @ nv - P[0] @ nv <= 1e-9 * size).all()
        area += a
        vol += a * np.dot(nv, P[0]) / 3.0
    assert abs(vol - hull.volume) <= 1e-10 * hull.volume
    assert abs(area - hull.area) <= 1e-10 * hull.area


def test_example04_sphere_builds_hull_segments():
    rad.FldLenRndSw("off")
    s = example04_sphere()
    items = list(O.walk(rad._reg, s))
    assert len(items) == 15
    # apex cones (1 + 15 faces), frusta (15 + 15 + 2 faces)
    assert [len(it[0].faces) for it in items] == [16] + [17] * 13 + [16]
    for it in items:
        _check_hull(it[0])


def test_example04_interior_field_is_two_thirds_of_M():
    """Example04's claim: inside the uniformly magnetized sphere B = 2/3 M
    (here M = 1 T along x), up to the polyhedral approximation."""
    rad.FldLenRndSw("off")
    s = example04_sphere()
    P, items = models.packed(s)
    OR.set_randomization(False)
    pts = np.array([[1e-3, 2e-3, 3e-3], [0.0, 0.31, 0.013], [0.1, -0.5, 0.2], [0.0, -0.6, 0.011]])
    B, _H = OR.field(OR.prepare_model(P), P["M"], pts)
    assert np.abs(B[:, 0] - 2.0 / 3.0).max() <= 2e-3
    assert np.abs(B[:2, 1:]).max() <= 1e-3


@pytest.mark.parametrize("pa,pb", [
    ([[0, 0], [2, 0], [2, 1], [0, 1]], [[0.5, 0.2], [1.5, 0.3], [1.0, 0.8]]),   # 4 -> 3
    ([[0, 0], [2, 0], [2, 1], [0, 1]], [[0, 0], [2, 0], [2, 1], [0, 1]]),       # prism, quads
    ([[0, 0], [1, 0], [0, 1]], [[0.3, 0.3]]),                                   # cone to an apex
    ([[0, 0], [1, 0]], [[0.5, -0.5], [0.5, 0.5]]),                               # two ridges
    ([[0, 0], [3, 0], [3, 2], [1, 3], [-1, 1]], [[1, 1], [2, 1]]),               # pentagon -> ridge
    ([[1, 0], [0, 1], [-1, 0], [0, -1]], [[0, 0], [2, 0], [2, 2], [0, 2]]),      # rotated squares
])
def test_layer_pairs_are_convex_hulls(pa, pb):
    rad.UtiDelAll()
    k = rad.ObjMltExtPgn([[pa, 0.0], [pb, 1.5]], [0, 0, 1])
    _check_hull(rad._reg[k])


def test_clockwise_layers_and_descending_z():
    sq = [[0, 0], [0, 1], [1, 1], [1, 0]]  # clockwise
    k = rad.ObjMltExtPgn([[sq, 2.0], [[[0.5, 0.5]], 1.0], [sq, 0.0]], [0, 0, 1])
    items = list(O.walk(rad._reg, k))
    assert len(items) == 2
    for it in items:
        _check_hull(it[0])


def test_improper_layers_raise_the_reference_errors():
    sq = [[0, 0], [1, 0], [1, 1], [0, 1]]
    with pytest.raises(RuntimeError, match="Radia::Error111"):
        rad.ObjMltExtPgn([[[[0, 0], [2, 0], [1, 0.3], [1, 2]], 0.0], [sq, 1.0]], [0, 0, 1])
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[sq, 0.0], [sq, 1.0], [sq, 0.5]], [0, 0, 1])       # z not monotonic
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[sq, 0.0], [sq, 0.0]], [0, 0, 1])                  # equal z
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[[[0, 0]], 0.0], [[[1, 0]], 1.0]], [0, 0, 1])      # no layer > 1 point
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[[[0, 0]], 0.0], [[[0, 0], [1, 0]], 1.0]], [0, 0, 1])   # apex + ridge
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[[[0, 0], [1, 0]], 0.0], [[[0, 1], [1, 1]], 1.0]], [0, 0, 1])  # parallel ridges
    with pytest.raises(RuntimeError, match="Radia::Error063"):
        rad.ObjMltExtPgn([[sq, 0.0]], [0, 0, 1])                              # one layer
