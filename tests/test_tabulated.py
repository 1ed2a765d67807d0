"""Tabulated isotropic materials (MatSatIsoTab): node tables, node
derivatives, the monotone correction and the cubic Hermite law, on the CPU.

The oracle (oracle/radora.c ora_tab_prepare, ora_tab_eval) and the product
tables (radiacuda_amd/objects.py tab_nodes, tab_eval) are two restatements of
radmater.cpp:341-646; both are pinned here against what the mathematics fixes:
3- and 5-point interpolant derivatives are exact on quadratic and quartic data,
a derivative below zero on increasing data is clamped to 0, the Hermite law
reproduces node values and slopes, a corrected segment's derivative touches
zero (discriminant 0), and the reference's index pattern for 4 to 7 nodes.
"""
import numpy as np
import pytest

from oracle import oracle as OR
from radiacuda_amd import _native as N
from radiacuda_amd import objects as O
from radiacuda_amd import rad
from tests import models


def _nodes(x, y):
    return [[float(a), float(b)] for a, b in zip(x, y)]


def _both(points):
    xp, yp, dp = O.tab_nodes(points)
    xo, yo, do = OR.tab_prepare(points)
    assert np.array_equal(xp, xo) and np.array_equal(yp, yo)
    scale = max(1.0, np.abs(dp).max())
    # closed-form (oracle) vs generic Lagrange (product) derivatives on
    # clustered nodes: rounding of the 5-point forms
    assert np.abs(dp - do).max() <= 1e-11 * scale
    return xp, yp, dp


def _grid(L, seed, top=3.0):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0.05, top, L - 1))
    return np.concatenate([[0.0], x])  # first point at the origin: no (0, 0) prepended


def test_quadratic_data_exact_derivatives():
    """Every 3- and 5-point derivative is exact on a quadratic; the last node
    takes a secant slope, placed as the reference places it (for 4 and 5 nodes
    it is the slope of an earlier segment, and with 5 nodes node 3 takes the
    derivative at node 2)."""
    p = lambda x: 0.7 * x + 0.2 * x * x
    dp = lambda x: 0.7 + 0.4 * x
    for L in (4, 5, 6, 7, 8, 9, 12, 20):
        x = _grid(L, L)
        y = p(x)
        xs, ys, d = _both(_nodes(x, y))
        assert len(xs) == L
        sec = lambda i: (y[i + 1] - y[i]) / (x[i + 1] - x[i])
        want = dp(x)
        if L == 4:
            want[3] = sec(1)
        elif L == 5:
            want[3] = dp(x[2])
            want[4] = sec(2)
        else:
            want[L - 1] = sec(L - 2)
        assert np.abs(d - want).max() <= 1e-11, (L, d - want)


def test_quartic_data_exact_five_point_derivatives():
    """The 5-point derivatives (nodes 3 .. L-4) are exact on a quartic."""
    p = lambda x: 0.5 * x + 0.1 * x ** 2 + 0.03 * x ** 3 + 0.01 * x ** 4
    dp = lambda x: 0.5 + 0.2 * x + 0.09 * x ** 2 + 0.04 * x ** 3
    for L in (6, 7, 8, 11, 16):
        x = _grid(L, 100 + L)
        y = p(x)
        _xs, _ys, d = _both(_nodes(x, y))
        k = np.arange(3, max(L - 3, 4))  # 5-point nodes: 3 .. L-4 (node 3 when L = 6)
        assert np.abs(d[k] - dp(x[k])).max() <= 1e-10 * np.abs(dp(x[k])).max(), L


def test_negative_derivative_on_increasing_data_is_clamped():
    """A knee: the quadratic through the first three (increasing) points falls
    at the origin; the reference sets that derivative to 0."""
    pts = [[0.0, 0.0], [1.0, 1.0], [1.1, 2.0], [3.0, 2.1], [5.0, 2.2], [8.0, 2.3]]
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    c = np.polyfit(x[:3], y[:3], 2)
    assert np.polyval(np.polyder(c), 0.0) < 0
    _xs, _ys, d = _both(pts)
    assert d[0] == 0.0


def test_origin_prepended_and_short_tables_refused():
    pts = [[0.1, 0.5], [0.2, 0.9], [0.4, 1.2], [1.0, 1.5]]
    x, y, _d = _both(pts)
    assert x[0] == 0.0 and y[0] == 0.0 and len(x) == 5
    with pytest.raises(ValueError):
        O.tab_nodes([[0.1, 0.5], [0.2, 0.9]])
    with pytest.raises(ValueError):
        OR.tab_prepare([[0.1, 0.5], [0.2, 0.9]])
    with pytest.raises(RuntimeError, match="Error071"):
        rad.MatSatIsoTab([[0.2, 0.5], [0.1, 0.9], [0.4, 1.2], [1.0, 1.5]])


def _hermite_coeffs(x, y, d, i):
    h = x[i + 1] - x[i]
    fm = y[i] - y[i + 1]
    a1 = d[i]
    a2 = -(3 * fm + (2 * d[i] + d[i + 1]) * h) / (h * h)
    a3 = (2 * fm + (d[i] + d[i + 1]) * h) / h ** 3
    return h, a1, a2, a3


def test_correction_makes_the_segment_derivative_touch_zero():
    """Data whose right-end slope is far above the secant: the Hermite segment
    overshoots, and the corrected left derivative is the one for which the
    segment's derivative has a double root (discriminant a2^2 - 3 a1 a3 = 0)."""
    pts = [[0.0, 0.0], [1.0, 0.1], [2.0, 0.2], [2.2, 0.9], [2.3, 1.9], [2.4, 3.0], [4.0, 3.1]]
    x, y, d = _both(pts)
    touched = []
    for i in range(len(x) - 1):
        h, a1, a2, a3 = _hermite_coeffs(x, y, d, i)
        disc = a2 * a2 - 3 * a1 * a3
        scale = a2 * a2 + abs(3 * a1 * a3)
        if a3 != 0.0 and abs(disc) <= 1e-9 * scale:
            touched.append(i)
    assert touched, d
    # a touched segment's derivative has its double root inside the segment
    # and stays >= 0: the law rises monotonically across it
    for i in touched:
        h, a1, a2, a3 = _hermite_coeffs(x, y, d, i)
        u = np.linspace(0.0, h, 4001)
        assert (a1 + 2 * a2 * u + 3 * a3 * u * u).min() >= -1e-9 * abs(a2) * h
        assert 0.0 < -a2 / (3 * a3) < h


def test_hermite_law_reproduces_nodes_and_slopes():
    pts = models.steel_points()
    x, y, d = _both(pts)
    t = np.stack([x, y, d], 1).ravel()
    for k in range(len(x)):
        assert O.tab_eval(x, y, d, x[k]) == pytest.approx(y[k], rel=1e-14, abs=1e-15)
        assert OR.tab_eval(t, x[k]) == pytest.approx(y[k], rel=1e-14, abs=1e-15)
    # one-sided slopes at interior nodes from both neighbouring segments
    for k in range(1, len(x) - 1):
        for i in (k - 1, k):
            h, a1, a2, a3 = _hermite_coeffs(x, y, d, i)
            u = x[k] - x[i]
            assert a1 + 2 * a2 * u + 3 * a3 * u * u == pytest.approx(d[k], rel=1e-9, abs=1e-9)
    # below the first node: y[0]; beyond the last: the last slope
    assert O.tab_eval(x, y, d, -1.0) == y[0]
    far = x[-1] + 2.5
    assert O.tab_eval(x, y, d, far) == pytest.approx(y[-1] + 2.5 * d[-1], rel=1e-15)
    rng = np.random.default_rng(5)
    for h in np.concatenate([rng.uniform(0, x[-1] * 1.5, 400), x * (1 + 1e-9)]):
        assert O.tab_eval(x, y, d, h) == pytest.approx(OR.tab_eval(t, h), rel=1e-14, abs=1e-16)


def test_example_steel_law_is_monotone():
    """The example's steel: every Hermite segment, after the correction, is
    non-decreasing (sampled finely), as the correction intends."""
    x, y, d = _both(models.steel_points())
    for i in range(len(x) - 1):
        u = np.linspace(x[i], x[i + 1], 2001)
        f = np.array([O.tab_eval(x, y, d, v) for v in u])
        assert np.diff(f).min() >= -1e-12 * max(1.0, abs(y[i + 1])), i


def test_material_records_and_mvsh():
    """Kind-3 records index one shared table per law; MatMvsH and the
    oracle's M(H) agree with the law."""
    top, steel = models.hybrid_undulator(n_periods=1)
    _P, items = models.packed(top)
    mats, tab = O.pack_materials(items)
    k3 = mats[:, 0] == 3
    assert k3.any() and (mats[k3, 1] == 0).all()
    L = int(mats[k3][0, 2])
    assert len(tab) == 3 * L
    x, y, d = O.tab_nodes(models.steel_points())
    assert np.array_equal(tab, np.stack([x, y, d], 1).ravel())
    assert items[0][0].mat.kind == O.MAT_NONLIN_TAB  # the half pole comes first
    OR.set_tables(tab)
    try:
        for hv in ([0.01, 0, 0], [0, 3e-4, 0], [0.2, -0.1, 0.05], [0, 0, 0]):
            H = np.array(hv, dtype=float)
            h = np.linalg.norm(H)
            Mw = (O.tab_eval(x, y, d, h) / h) * H if h > 0 else np.zeros(3)
            Mo = OR.mat_M(mats[k3][0], H)
            Mn = N.mat_M(mats[k3][0], H, tab)
            assert np.abs(Mo - Mw).max() <= 1e-14
            assert np.abs(Mn - Mw).max() <= 1e-14
            assert np.abs(np.array(rad.MatMvsH(top, "m", list(H))) - Mw).max() <= 1e-14
    finally:
        OR.set_tables(None)


def test_oracle_relaxation_reaches_the_newton_fixed_point():
    """The oracle's method-4 relaxation with the tabulated law converges to the
    FP64 Newton fixed point of the same (FP64) system to FP32 dot-product
    accuracy: the test-side reference the device solve is held to."""
    OR.set_randomization(False)
    top, _steel = models.hybrid_undulator(n_periods=1)
    P, items = models.packed(top)
    mats, tab = O.pack_materials(items)
    n = len(items)
    hext = np.zeros((n, 3))
    Qo, _mask = models.oracle_assemble(P)
    OR.set_tables(tab)
    try:
        Mo, _H, iters, _stats = OR.relax_m4(Qo.astype(np.float32), mats, hext, 1e-8, 20000)
    finally:
        OR.set_tables(None)
    Mn = models.fp64_fixed_point(Qo, mats, hext, M0=Mo, tab=tab)
    scale = np.abs(Mn).max()
    assert iters < 20000
    assert np.abs(Mo - Mn).max() <= 1e-4 * scale
    # the fixed point: M = M_mat(Q M + hext) to rounding
    H = (Qo @ Mn.ravel()).reshape(n, 3) + hext
    F = np.stack([models.mat_M_and_D(mats[e], H[e], tab)[0] for e in range(n)])
    assert np.abs(F - Mn).max() <= 1e-12 * scale
