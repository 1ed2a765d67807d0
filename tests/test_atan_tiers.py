"""The short atan polynomials that the field sweeps use when every lane of a
wave has |t| < tan(pi/16) (csrc/rh_math.hpp, rh_atan2_wave).  The
coefficients are read from the header and evaluated the way the kernels
evaluate them (Horner in t^2, then t + t^3 p), in float64 and float32, against
numpy's arctan over the whole range of the tier."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "radiacuda_amd", "csrc",
                   "rh_math.hpp")


def _source():
    with open(HDR) as fh:
        return fh.read()


def test_tier_bound_is_tan_pi_16():
    src = _source()
    k64 = float(re.search(r"kTanPi16\s*=\s*([0-9.eE+-]+)", src).group(1))
    assert abs(k64 - np.tan(np.pi / 16)) <= 1e-16
    k32 = float(re.search(r"fabsf\(y\)\s*<\s*([0-9.]+)f\s*\*\s*x", src).group(1))
    assert abs(k32 - np.tan(np.pi / 16)) <= 1e-8


def test_fp64_degree6_polynomial_within_bound():
    src = _source()
    body = re.search(r"d_atan16_p\[7\]\s*=\s*\{([^}]*)\}", src).group(1)
    c = np.array([float(v) for v in body.replace("\n", " ").split(",")])
    assert len(c) == 7
    t = np.linspace(-np.tan(np.pi / 16), np.tan(np.pi / 16), 200001)
    u = t * t
    p = np.full_like(t, c[6])
    for i in range(5, -1, -1):
        p = p * u + c[i]
    err = np.abs((t + t * u * p) - np.arctan(t))
    assert err.max() <= 2e-17 + 4e-16 * np.abs(np.arctan(t)).max()
    # relative to |atan t| away from 0 (the kernels sum solid angles)
    nz = np.abs(t) > 1e-3
    assert (err[nz] / np.abs(np.arctan(t[nz]))).max() <= 5e-16


def _mp_fit_error(c, hi, count=2001):
    """max |t + t^3 p(t^2) - atan t| over [0, hi] in 40-digit arithmetic (the
    polynomial is odd), with the coefficients as stored."""
    import mpmath as mp
    mp.mp.dps = 40
    cc = [mp.mpf(float(v)) for v in c]
    worst = mp.mpf(0)
    for i in range(count):
        t = mp.mpf(hi) * i / (count - 1)
        u = t * t
        p = cc[-1]
        for v in reversed(cc[:-1]):
            p = p * u + v
        worst = max(worst, abs(t + t * u * p - mp.atan(t)))
    return float(worst)


def test_fp64_degree6_fit_error_in_exact_arithmetic():
    """The fit itself, free of float64 rounding: within 1e-17 of atan on the
    whole tier (7.4e-18 documented).  numpy's arctan carries about 1 ulp of its
    own error, so only exact arithmetic resolves this claim."""
    src = _source()
    body = re.search(r"d_atan16_p\[7\]\s*=\s*\{([^}]*)\}", src).group(1)
    c = [float(v) for v in body.replace("\n", " ").split(",")]
    assert _mp_fit_error(c, np.tan(np.pi / 16)) <= 1e-17


def test_fp32_degree2_fit_error_in_exact_arithmetic():
    src = _source()
    m = re.search(r"const float p = ([-0-9.]+)f \+ u \* \(([-0-9.]+)f \+ u \* ([-0-9.]+)f\);", src)
    c = [float(np.float32(m.group(i))) for i in (1, 2, 3)]
    assert _mp_fit_error(c, float(np.float32(0.19891237))) <= 2e-9


def test_fp32_degree2_polynomial_within_bound():
    src = _source()
    m = re.search(r"const float p = ([-0-9.]+)f \+ u \* \(([-0-9.]+)f \+ u \* ([-0-9.]+)f\);", src)
    c = [np.float32(m.group(i)) for i in (1, 2, 3)]
    t = np.linspace(-np.tan(np.pi / 16), np.tan(np.pi / 16), 200001).astype(np.float32)
    u = t * t
    p = c[0] + u * (c[1] + u * c[2])
    r = t + (t * u) * p
    assert r.dtype == np.float32
    err = np.abs(r.astype(np.float64) - np.arctan(t.astype(np.float64)))
    assert err.max() <= 3e-8
    nz = np.abs(t) > 1e-3
    assert (err[nz] / np.abs(np.arctan(t[nz].astype(np.float64)))).max() <= 2e-7
