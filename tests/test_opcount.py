"""The FLOP counts behind the benchmark's VALU rooflines, counted from the
kernel math itself: the host build instantiates the FP64 integrals with an
op-counting scalar (radiacuda_amd/csrc/rh_opcount.hpp, rhc_opcount in
_rhcheck.so) and counts every operation on the branches the points take.

Weights (fixed in advance of any measurement): + - x / sqrt count 1, a fused
multiply-add 2, a log 20, an atan2 30.  The counting scalar sees the
arithmetic done in the integrals' own precision; the few FP64 operations the
kernels add around them are counted here by hand and listed next to each
figure (vertex offsets formed in FP64 before rounding, the weighted
accumulation of each term, the 1/4pi scale)."""
import ctypes
import os

import numpy as np
import pytest

import bench
from radiacuda_amd import _abi
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from radiacuda_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "radiacuda_amd", "_rhcheck.so")
WEIGHTS = np.array([1, 1, 1, 1, 2, 20, 30])  # add, mul, div, sqrt, fma, log, atan2

# FP64 operations outside the templated integrals, per evaluation
EXTRA_MESH_EDGE = 6 + 6          # two vertex offsets; 3 FMA of L * w
EXTRA_MESH_TRI = 9 + 1 + 6       # three vertex offsets; 2 * atan2; 3 FMA of Omega * g
EXTRA_TET_BLOCK = 12 + 72 + 6 + 1 + 72 + 6   # offsets; 6 edges x 6 FMA; 2*atan2 and the running
#                                             sum (3 faces); 4 pi - sum; 4 faces x 6 x (2 mul + sub); 1/4pi
EXTRA_TET_FIELD = 12 + 36 + 6 + 1 + 24 + 3   # offsets; 6 edges x 3 FMA; 2*atan2 and sum; 4 pi - sum;
#                                             4 faces x 3 x (mul + sub); 1/4pi


@pytest.fixture(scope="module")
def counts():
    if not os.path.exists(LIB):
        from radiacuda_amd import build
        build.build()
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    rad.UtiDelAll()
    _top, iron, _coil, info = W.c_yoke()
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    desc, keep = _abi.model_desc(P)
    rng = np.random.default_rng(0)
    M = np.ascontiguousarray(rng.normal(size=(len(items), 3)))
    pts = np.ascontiguousarray(W.gap_grid(info, 100)[rng.choice(1_000_000, 8, replace=False)])
    out = np.zeros(28)
    assert L.rhc_opcount(ctypes.byref(desc), M.ctypes.data_as(dp), pts.ctypes.data_as(dp),
                         ctypes.c_int64(len(pts)), out.ctypes.data_as(dp)) == 0
    rad.UtiDelAll()
    return out.reshape(4, 7)


def test_transcendentals_per_term(counts):
    # one log per edge term, one atan2 per triangle, 6 logs and 3 atan2 per
    # tetrahedron (the fourth solid angle from the sum rule)
    assert counts[0, 5] == 1 and counts[0, 6] == 0
    assert counts[1, 5] == 0 and counts[1, 6] == 1
    assert counts[2, 5] == 6 and counts[2, 6] == 3
    assert counts[3, 5] == 6 and counts[3, 6] == 3
    assert counts[0, 3] == 2 and counts[1, 3] == 3 and counts[2, 3] == 4


def test_bench_flop_constants_are_the_counted_ones(counts):
    f = counts @ WEIGHTS
    assert round(f[0] + EXTRA_MESH_EDGE) == bench.FLOP_MESH_EDGE
    assert round(f[1] + EXTRA_MESH_TRI) == bench.FLOP_MESH_TRI
    assert round(f[2] + EXTRA_TET_BLOCK) == bench.FLOP_TET_BLOCK
    assert round(f[3] + EXTRA_TET_FIELD) == bench.FLOP_TET_FIELD
