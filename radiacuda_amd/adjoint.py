"""Differentiable field and linear-material relaxation (torch.autograd).

Forward passes run the device kernels behind the C ABI; backward passes run
the adjoint kernels (rh_field_adjoint_dev, rh_relax_adjoint_dev).  Geometry
enters through polyhedron vertex coordinates: every forward call repacks the
elements for the given vertices and the kernels recompute the interaction
blocks and face integrals from them, so no derivative of the matrix is ever
stored.

    geom = Geometry(packed)                      # objects.pack_elements output
    B = field(M, verts, pts, geom)               # [np, 3], differentiable in M, verts, pts
    M = relax_linear(Mr, verts, geom, material)  # solves (I - K Q(verts)) M = Mr (+ K Hext)

Tensors are float64 on the GPU (torch's device memory and current stream);
the math runs in the HIP kernels.  There is no CPU fallback.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from . import _native as N

_vp = ctypes.c_void_p


def _dev(t):
    return t.detach().to(device="cuda", dtype=torch.float64).contiguous()


def _ptr(t):
    return _vp(t.data_ptr()) if t is not None else None


class Geometry:
    """Fixed element topology (faces, transforms, kinds) whose polyhedron
    vertex coordinates vary; ``model(verts)`` builds the device model."""

    def __init__(self, packed):
        self.packed = packed
        self.n = len(packed["type"])
        self.nvert = len(packed["verts"])
        self.vert_off = np.asarray(packed["vert_off"])
        self.identity_transforms = bool(
            np.all(np.asarray(packed["copy_off"])[1:] - np.asarray(packed["copy_off"])[:-1] == 1)
            and np.allclose(packed["cM"], np.eye(3)) and np.allclose(packed["cV"], 0)
            and np.allclose(packed["cS"], 1))

    def model(self, verts):
        P = dict(self.packed)
        P["verts"] = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
        return N.Model(P, stream=N.stream_handle())


def _field_dev(model, M, pts, field):
    out = torch.empty_like(pts)
    L = N.lib()
    b = out if field == "b" else None
    h = out if field == "h" else None
    N.check(L.rh_field_dev(model.h, _ptr(M), _ptr(pts), pts.shape[0], _ptr(b), _ptr(h), N.RH_FP64))
    return out


def _field_adjoint_dev(model, M, pts, g, field, want_m, want_v, want_x, nvert):
    L = N.lib()
    gM = torch.zeros((model.n, 3), dtype=torch.float64, device="cuda") if want_m else None
    gV = torch.zeros((nvert, 3), dtype=torch.float64, device="cuda") if want_v else None
    gX = torch.zeros_like(pts) if want_x else None
    N.check(L.rh_field_adjoint_dev(model.h, _ptr(M), _ptr(pts), pts.shape[0], _ptr(g),
                                   0 if field == "b" else 1, _ptr(gM), _ptr(gV), _ptr(gX)))
    return gM, gV, gX


class _FieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, verts, pts, geom, field):
        Md, Vd, Pd = _dev(M).reshape(-1, 3), _dev(verts).reshape(-1, 3), _dev(pts).reshape(-1, 3)
        model = geom.model(Vd.cpu().numpy())
        out = _field_dev(model, Md, Pd, field)
        ctx.save_for_backward(Md, Vd, Pd)
        ctx.geom, ctx.field, ctx.model = geom, field, model
        return out.to(M.device)

    @staticmethod
    def backward(ctx, gout):
        Md, Vd, Pd = ctx.saved_tensors
        need_m, need_v, need_x = ctx.needs_input_grad[:3]
        gM, gV, gX = _field_adjoint_dev(ctx.model, Md, Pd, _dev(gout), ctx.field, need_m, need_v,
                                        need_x, ctx.geom.nvert)
        ctx.model.close()
        return gM, gV, gX, None, None


def field(M, verts, pts, geom, field="b"):
    """B (field='b') or H ('h') of elements with magnetization M [n,3] (element
    frames) and polyhedron vertices verts [nvert,3] at pts [np,3]."""
    return _FieldFn.apply(M, verts, pts, geom, field)


class _RelaxFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Mr, verts, geom, ktensor, hext, prec, max_iter):
        if not geom.identity_transforms:
            raise N.RadHipError("relax_linear differentiates models without transform copies only")
        Vd = _dev(verts).reshape(-1, 3)
        model = geom.model(Vd.cpu().numpy())
        n = geom.n
        mats = torch.zeros((n, 16), dtype=torch.float64, device="cuda")
        mats[:, 0] = 1.0
        mats[:, 4:13] = _dev(ktensor).reshape(n, 9)
        mats[:, 13:16] = _dev(Mr).reshape(n, 3)
        hx = torch.zeros((n, 3), dtype=torch.float64, device="cuda") if hext is None else _dev(hext)
        Q = model.assemble(N.RH_FP64)
        M = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        H = torch.zeros_like(M)
        opts = _abi.RelaxOpts(float(prec), int(max_iter), 0, 0.0, 9, -1)
        res = _abi.RelaxResult()
        L = N.lib()
        N.check(L.rh_relax_dev(Q.h, _ptr(mats), _ptr(hx), _ptr(M), _ptr(H), ctypes.byref(opts),
                               ctypes.byref(res)))
        if not res.converged:
            raise N.RadHipError("relaxation did not reach prec %g (misfit %g)" % (prec, res.misfit))
        ctx.save_for_backward(Vd, M, mats)
        ctx.geom, ctx.model, ctx.Q, ctx.prec, ctx.max_iter = geom, model, Q, prec, max_iter
        ctx.hext = hx
        return M.to(Mr.device)

    @staticmethod
    def backward(ctx, gM):
        Vd, M, mats = ctx.saved_tensors
        geom, model, Q = ctx.geom, ctx.model, ctx.Q
        n = geom.n
        L = N.lib()
        lam = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        opts = _abi.RelaxOpts(float(ctx.prec), int(ctx.max_iter), 0, 0.0, 9, 0)
        res = _abi.RelaxResult()
        N.check(L.rh_relax_adjoint_dev(Q.h, _ptr(mats), _ptr(_dev(gM)), _ptr(lam), ctypes.byref(opts),
                                       ctypes.byref(res)))
        K = mats[:, 4:13].reshape(n, 3, 3)
        mu = torch.einsum("nba,nb->na", K, lam)  # K^T lam
        gMr = lam
        gK = None
        if ctx.needs_input_grad[3]:
            # dL/dK_ab = lam_a * H_b with H = Q M + Hext at the collocation points
            H = torch.zeros_like(M)
            X = torch.as_tensor(model.colloc(), dtype=torch.float64, device="cuda")
            N.check(L.rh_field_dev(model.h, _ptr(M), _ptr(X), n, None, _ptr(H), N.RH_FP64))
            H = H + ctx.hext
            gK = torch.einsum("na,nb->nab", lam, H).reshape(n, 9)
        gV = None
        ghext = mu if ctx.needs_input_grad[4] else None
        if ctx.needs_input_grad[1]:
            # lam^T K dQ/dV M: the field at the collocation points (H, element
            # frames) differentiated in the source vertices and in the points,
            # which move with their element's vertex mean
            X = torch.as_tensor(model.colloc(), dtype=torch.float64, device="cuda")
            _g, gVs, gX = _field_adjoint_dev(model, M, X, mu.contiguous(), "h", False, True, True,
                                             geom.nvert)
            gV = gVs
            vo = geom.vert_off
            counts = torch.as_tensor(np.diff(vo), dtype=torch.float64, device="cuda")
            owner = torch.as_tensor(np.repeat(np.arange(n), np.diff(vo)), device="cuda")
            gV = gV + (gX / counts[:, None])[owner]
        Q.close()
        model.close()
        return gMr, gV, None, gK, ghext, None, None


def relax_linear(Mr, verts, geom, ktensor, hext=None, prec=1e-13, max_iter=2000):
    """Magnetization M [n,3] of linear-material elements: (I - K Q) M = K Hext + Mr,
    K = ktensor [n,9] (row-major susceptibility tensors, element frames),
    Mr [n,3] remanent magnetization, Q assembled in FP64 from verts."""
    return _RelaxFn.apply(Mr, verts, geom, ktensor, hext, prec, max_iter)
