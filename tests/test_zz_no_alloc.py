"""rh_field_dev_noalloc (the field entry point that allocates nothing) against
rh_field_dev on the same device arrays, through the C ABI.

Bar: FP64 <= 1e-10 of the field scale (the reduced element splits change the
summation grouping; with the same splits the results are identical), FP32
integrals <= 1e-6.  Two calls in a row, the second with fewer points, so the
second call runs over the first call's leftover scratch contents.
"""
import ctypes

import numpy as np
import pytest

from radiacuda_amd import _abi
from radiacuda_amd import _native as N
from radiacuda_amd import rad
from tests import models

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _model():
    # tetrahedra and blocks: two element kinds, so the sweep needs partial sums
    rng = np.random.default_rng(3)
    ids = models.tet_mesh((3, 2, 2), 5.0, origin=(-7, -5, -5), jitter=0.1, seed=4)
    ids.append(rad.ObjRecMag([14, 2, 1], [4, 6, 5], list(rng.uniform(-1, 1, 3))))
    ids.append(rad.ObjRecMag([-14, -3, 2], [5, 3, 4], list(rng.uniform(-1, 1, 3))))
    return rad.ObjCnt(ids)


def _call(L, fn, m, dM, dP, n, dB, dH, prec):
    st = fn(m.h, _vp(dM.data_ptr()), _vp(dP.data_ptr()), n, _vp(dB.data_ptr()), _vp(dH.data_ptr()), prec)
    return st


def test_noalloc_field_matches_rh_field_dev():
    import torch
    L = N.lib()
    top = _model()
    P, _items = models.packed(top)
    rng = np.random.default_rng(11)
    dev = torch.device("cuda:0")
    f64 = dict(dtype=torch.float64, device=dev)
    with N.Model(P) as m:
        dM = torch.tensor(np.asarray(P["M"]).reshape(-1, 3), **f64)
        # 500,000 points: the occupancy split would need 72 MB of partials,
        # so the no-allocation call runs with fewer splits; 3,000 points: many
        # splits, all within the fixed scratch
        for npts in (500_000, 3_000):
            print("case: noalloc field, %d points" % npts, flush=True)
            pts = torch.tensor(rng.uniform(-25, 25, (npts, 3)), **f64)
            for prec, tol in ((N.RH_FP64, 1e-10), (N.RH_MIXED, 1e-6)):
                B0 = torch.zeros((npts, 3), **f64)
                H0 = torch.zeros((npts, 3), **f64)
                B1 = torch.full((npts, 3), np.nan, **f64)
                H1 = torch.full((npts, 3), np.nan, **f64)
                N.check(_call(L, L.rh_field_dev, m, dM, pts, npts, B0, H0, prec))
                N.check(_call(L, L.rh_field_dev_noalloc, m, dM, pts, npts, B1, H1, prec))
                b0, b1 = B0.cpu().numpy(), B1.cpu().numpy()
                h0, h1 = H0.cpu().numpy(), H1.cpu().numpy()
                assert np.isfinite(b1).all() and np.isfinite(h1).all()
                assert np.abs(b1 - b0).max() <= tol * np.abs(b0).max()
                assert np.abs(h1 - h0).max() <= tol * np.abs(h0).max()


def test_noalloc_field_refuses_beyond_scratch():
    import torch
    L = N.lib()
    top = _model()
    P, _items = models.packed(top)
    dev = torch.device("cuda:0")
    f64 = dict(dtype=torch.float64, device=dev)
    with N.Model(P) as m:
        npts = 700_000  # two element kinds: at most 699,050 points fit
        print("case: noalloc field refusal, %d points" % npts, flush=True)
        dM = torch.tensor(np.asarray(P["M"]).reshape(-1, 3), **f64)
        pts = torch.zeros((npts, 3), **f64)
        B = torch.zeros((npts, 3), **f64)
        st = L.rh_field_dev_noalloc(m.h, _vp(dM.data_ptr()), _vp(pts.data_ptr()), npts,
                                    _vp(B.data_ptr()), None, N.RH_FP64)
        assert st == _abi.RH_ERR_ARG
        assert b"699050" in L.rh_last_error()
