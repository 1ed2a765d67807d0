#!/usr/bin/env python3
"""Benchmark of the three paths on one MI355X: interaction-matrix assembly,
relaxation solve and field evaluation, FP64 against mixed precision, with
every input already resident in HBM.

Workloads (seeded, synthetic, built through the drop-in API; no files):

* large yoke (the headline, `value`): 40,000 jittered tetrahedra of a C-shaped
  dipole yoke, MatSatIsoFrm steel, two racetrack coils; a step = assemble the
  120,000 x 120,000 interaction matrix + coil field at the collocation points
  + relax to prec 1e-5 from the remanent state.  FP64 and mixed (FP32 matrix,
  FP64 self blocks, FP64 refinement); the matrix is stored in the
  symmetric-block layout (5 planes per element row: 64 GB in FP64, 32 GB in
  FP32) unless --dense-rows asks for row-major 3x3 blocks (115 / 57.6 GB);
* Halbach ring: 20,000 hexahedral polyhedra, MatLin 0.06/0.17, B at 2,000,000
  points in the bore; a step = one field sweep, FP64 and FP32 integrals;
* C-yoke: 8,000 tetrahedra; a step = assemble + relax (prec 1e-4) + B of the
  iron and the coils on a 100^3 gap grid, FP64 and mixed.

The roofline is that of the headline's dominant kernel, the FP64 matrix pass
(HBM-bound).  The CPU baseline is the FP64 oracle on this machine's host
cores, on a bounded sample of the headline workload.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Q-blocks/s, rad.Solve s, Fld points/s: mixed vs FP64 GPU at stated error"
# Algorithmic FLOPs per evaluation of the FP64 integrals, counted from the
# kernel math instantiated with an op-counting scalar (tests/test_opcount.py
# recounts them): + - * / sqrt count 1, a fused multiply-add 2, a log 20 and
# an atan2 30 (fixed weights).  Per (point, tetrahedron) field 480 and block
# 669 (6 logs, 3 atan2, 4 sqrt); mesh form per (point, unique edge) 56 (1 log,
# 2 sqrt), per (point, unique triangle) 93 (1 atan2, 3 sqrt).
FLOP_TET_FIELD = 480
FLOP_TET_BLOCK = 669
FLOP_MESH_EDGE = 56
FLOP_MESH_TRI = 93
PEAK_FP64_VALU = 78.6   # TFLOP/s, MI355X vector FP64
PEAK_FP32_VALU = 157.3  # TFLOP/s, MI355X vector FP32
PEAK_HBM = 8000.0       # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Ctx:
    """Shared handles: torch device, the library, the current stream."""

    def __init__(self):
        import torch
        from radiacuda_amd import _native as N
        assert torch.cuda.is_available(), "no GPU visible"
        self.torch = torch
        self.N = N
        self.dev = torch.device("cuda:0")
        torch.cuda.init()
        self.stream = torch.cuda.current_stream()
        self.sh = ctypes.c_void_p(self.stream.cuda_stream)
        self.L = N.lib()

    def t64(self, a):
        return self.torch.tensor(np.ascontiguousarray(a), dtype=self.torch.float64, device=self.dev)

    def zeros(self, *shape):
        return self.torch.zeros(shape, dtype=self.torch.float64, device=self.dev)

    def sync(self):
        self.torch.cuda.synchronize()


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class Coils:
    """Packed current sources and a device-side evaluation into d_out."""

    def __init__(self, ctx, cur):
        self.ctx = ctx
        self.kind = np.ascontiguousarray(cur["kind"], dtype=np.int32)
        self.par = np.ascontiguousarray(cur["par"])
        self.tm, self.tv, self.ts = (np.ascontiguousarray(cur[k]) for k in ("cM", "cV", "cS"))
        self.cur = cur

    def add(self, d_points, count, d_out):
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        self.ctx.N.check(self.ctx.L.rh_current_field_dev(
            0, self.ctx.sh, len(self.kind), self.kind.ctypes.data_as(ip), self.par.ctypes.data_as(dp),
            self.tm.ctypes.data_as(dp), self.tv.ctypes.data_as(dp), self.ts.ctypes.data_as(dp),
            _vp(d_points), count, _vp(d_out)))


def timed_steps(ctx, step, steps, warmup, label):
    """warmup untimed steps, then `steps` timed ones bracketed by device
    synchronization; returns (wall seconds, per-phase seconds summed over the
    timed steps, per-family kernel statistics)."""
    N = ctx.N
    for w in range(warmup):
        step(None)
        log("%s: warmup %d done" % (label, w))
    phase = {}
    N.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        step(phase)
        log("%s: step %d done" % (label, s))
    ctx.sync()
    wall = time.perf_counter() - t0
    kt = {f: N.kernel_time(f) for f in ("asm", "field", "matvec", "refine")}
    N.set_timing(False)
    return wall, phase, kt


def _acc(phase, key, dt):
    if phase is not None:
        phase[key] = phase.get(key, 0.0) + dt


# ----------------------------------------------------------------- large yoke
def run_large_yoke(ctx, args, cpu):
    from radiacuda_amd import _abi
    from radiacuda_amd import objects as O
    from radiacuda_amd import rad
    from radiacuda_amd import workloads as W
    N = ctx.N
    log("== large yoke: building 40,000 tetrahedra")
    t0 = time.perf_counter()
    rad.UtiDelAll()
    top, iron, coil, info = W.c_yoke(ny=50)
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P, stream=ctx.sh)
    X = model.colloc()
    coils = Coils(ctx, O.pack_currents(list(O.walk(rad._reg, coil))))
    pts = W.gap_grid(info, 20)[::97]
    log("large yoke: %d elements, model built in %.1f s" % (n, time.perf_counter() - t0))
    d_X, d_mats = ctx.t64(X), ctx.t64(mats)
    d_hext, d_M, d_H = ctx.zeros(n, 3), ctx.zeros(n, 3), ctx.zeros(n, 3)
    d_pts, d_B = ctx.t64(pts), ctx.zeros(len(pts), 3)
    out = {}
    for label, prec in (("fp64", N.RH_FP64), ("mixed", N.RH_MIXED)):
        if args.only and label != args.only:
            continue
        log("large yoke %s: allocating and assembling the matrix" % label)
        Q = model.assemble(prec, dense_rows=args.dense_rows)
        planes, mld = Q.format()
        opts = _abi.RelaxOpts(args.prec_large, args.max_iter, 0, 0.0, 9, 0)
        res = _abi.RelaxResult()

        def step(phase):
            t = time.perf_counter()
            Q.reassemble()
            t1 = time.perf_counter()
            d_hext.zero_()
            coils.add(d_X, n, d_hext)
            N.check(ctx.L.rh_relax_dev(Q.h, _vp(d_mats), _vp(d_hext), _vp(d_M), _vp(d_H),
                                       ctypes.byref(opts), ctypes.byref(res)))
            t2 = time.perf_counter()
            _acc(phase, "asm", t1 - t)
            _acc(phase, "solve", t2 - t1)

        wall, phase, kt = timed_steps(ctx, step, args.steps, args.warmup, "large yoke " + label)
        K = args.steps
        # field of the relaxed state at sampled gap points (for the errors)
        N.check(ctx.L.rh_field_dev(model.h, _vp(d_M), _vp(d_pts), len(pts), _vp(d_B), None, prec))
        ctx.sync()
        B = d_B.cpu().numpy()
        B32 = None
        if prec == N.RH_FP64:
            # the FP32 field integrals on the same (FP64) magnetization
            N.check(ctx.L.rh_field_dev(model.h, _vp(d_M), _vp(d_pts), len(pts), _vp(d_B), None, N.RH_MIXED))
            ctx.sync()
            B32 = d_B.cpu().numpy()
        r = dict(ms_per_step=1e3 * wall / K, asm_ms=1e3 * phase["asm"] / K,
                 solve_s=phase["solve"] / K, iters=res.iters, misfit=res.misfit,
                 converged=bool(res.converged), refined=res.refined,
                 kernel_ms={f: kt[f][0] for f in kt}, kernel_counts={f: kt[f][1] for f in kt},
                 M=d_M.cpu().numpy(), B=B, B32=B32, planes=planes, mld=mld)
        Q.close()
        out[label] = r
        log("large yoke %s: %.1f ms/step (asm %.1f ms, solve %.3f s in %d it, misfit %.2e)"
            % (label, r["ms_per_step"], r["asm_ms"], r["solve_s"], r["iters"], r["misfit"]))
    model.close()
    n3 = 3 * n
    sec = {"elements": n, "matrix": "%dx%d" % (n3, n3), "prec": args.prec_large}
    for label, r in out.items():
        es = 8.0 if label == "fp64" else 4.0
        mv = r["kernel_ms"]["matvec"]
        # stored matrix bytes (every element row, padding included) + the
        # iterate read and the product written
        stored = es * (r["planes"] * n * r["mld"] if r["planes"] else n3 * r["mld"])
        bytes_pass = stored + 72.0 * n
        r["bytes_pass"] = bytes_pass
        sec[label] = {
            "layout": ("symmetric blocks, %d planes of %d" % (r["planes"], r["mld"]) if r["planes"]
                       else "row-major 3x3 blocks, rows of %d" % r["mld"]),
            "step_ms": r["ms_per_step"], "end_to_end_qblocks_per_s": n * n / (r["ms_per_step"] * 1e-3),
            "asm_ms": r["asm_ms"], "asm_qblocks_per_s": n * n / (r["asm_ms"] * 1e-3),
            "asm_kernel_ms": r["kernel_ms"]["asm"], "solve_s": r["solve_s"], "solve_iters": r["iters"],
            "solve_misfit": r["misfit"], "matrix_passes": r["kernel_counts"]["matvec"] // max(args.steps, 1),
            "matvec_kernel_ms": mv, "matvec_TBps": bytes_pass / (mv * 1e-3) / 1e12 if mv > 0 else None,
            "matvec_hbm_frac": bytes_pass / (mv * 1e-3) / (PEAK_HBM * 1e9) if mv > 0 else None,
            "matrix_GB": stored / 1e9}
        if label == "mixed":
            sec[label]["fp64_refinement_passes"] = r["refined"]
            sec[label]["refine_field_ms"] = r["kernel_ms"]["refine"]
    if "fp64" in out and "mixed" in out:
        r64, rmx = out["fp64"], out["mixed"]
        sec["mixed"]["speedup_vs_fp64"] = r64["ms_per_step"] / rmx["ms_per_step"]
        sec["mixed"]["M_err_vs_fp64"] = float(np.abs(rmx["M"] - r64["M"]).max() / np.abs(r64["M"]).max())
        sec["mixed"]["B_err_vs_fp64"] = float(np.abs(rmx["B"] - r64["B"]).max() / np.abs(r64["B"]).max())
        sec["mixed"]["fp32_field_err_same_M"] = float(np.abs(r64["B32"] - r64["B"]).max() / np.abs(r64["B"]).max())
        sec["mixed"]["error_note"] = (
            "M_err and B_err compare two solves that each stop at misfit <= prec (%.0e) on different paths: "
            "at chi ~ 2000 the stopping tolerance, not the precision, sets their difference (the GPU tests "
            "take both on to prec 1e-9, where the refined mixed state's gap field is the FP64 one to 1e-6); "
            "fp32_field_err_same_M is the FP32 field integrals against the FP64 ones on one magnetization"
            % args.prec_large)
    if cpu and "fp64" in out:
        from oracle import oracle as OR
        OR.set_randomization(False)
        om = OR.prepare_model(P)
        Bo, _ = OR.field(om, out["fp64"]["M"].ravel(), pts)
        sec["fp64"]["B_err_vs_oracle_sampled"] = float(np.abs(out["fp64"]["B"] - Bo).max() / np.abs(Bo).max())
        if "mixed" in out:
            sec["mixed"]["B_err_vs_oracle_sampled"] = float(np.abs(out["mixed"]["B"] - Bo).max()
                                                            / np.abs(Bo).max())
        sec["error_note"] = ("B of the relaxed state at %d gap points vs the oracle's field of the same "
                             "FP64 magnetization; the oracle restates the reference's forms, which lose "
                             "accuracy on this mesh's 14 tetrahedra with a face within 1e-7 of -z and on "
                             "jittered faces (the GPU tests pin those elements and the whole model at 50 "
                             "digits: the device FP64 field is within 1e-14 of the 50-digit sums)" % len(pts))
    return sec, out, P, mats


def cpu_baseline_large(P, mats, n, r64, args):
    """The FP64 oracle on the host cores, on a bounded sample of the large
    yoke: assembly of sampled rows (extrapolated by row count) and the
    method-4 Gauss-Seidel solve of a 1,600-tetrahedron yoke of the same
    construction to prec 1e-5 (its own sweep count; the sweep cost scales as
    n^2).  Timed single-threaded (the reference's assembly and method 4 are
    serial) and with OpenMP over all cores."""
    from oracle import oracle as OR
    from radiacuda_amd import objects as O
    from radiacuda_amd import rad
    from radiacuda_amd import workloads as W
    OR.set_randomization(False)
    threads = len(os.sched_getaffinity(0))
    om = OR.prepare_model(P)
    rows = args.cpu_rows
    rng = np.random.default_rng(0)
    samp = np.sort(rng.choice(n, rows, replace=False))
    legs = {}
    # the shrunken model for the solve's sweep count
    rad.UtiDelAll()
    _top, iron, coil, _info = W.c_yoke(ny=2)
    items = list(O.walk(rad._reg, iron))
    Ps = O.pack_elements(items)
    ns = len(items)
    oms = OR.prepare_model(Ps)
    Xs = OR.colloc(oms)
    hext = _oracle_currents(OR, O.pack_currents(list(O.walk(rad._reg, coil))), Xs)
    mats_s = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    Qs = OR.assemble(oms).astype(np.float32)
    for label, th in (("single_thread", 1), ("openmp", threads)):
        OR.set_threads(th)
        tc = time.perf_counter()
        if th == 1:
            for i in samp:
                OR.assemble(om, rows=(int(i), int(i) + 1))
            nrows = rows
        else:
            # one contiguous block of rows per call, so every thread has rows
            r0 = int(samp[0]) if int(samp[0]) + rows * th <= n else n - rows * th
            OR.assemble(om, rows=(r0, r0 + rows * th))
            nrows = rows * th
        t_asm = (time.perf_counter() - tc) * n / nrows
        tc = time.perf_counter()
        _M, _H, it, stats = OR.relax_m4(Qs, mats_s, hext, args.prec_large, 3000)
        t_small = time.perf_counter() - tc
        t_solve = t_small / max(it, 1) * it * (n / ns) ** 2
        t_cpu = t_asm + t_solve
        legs[label] = {"value": n * n / t_cpu, "cores": th, "asm_s": t_asm, "solve_s": t_solve,
                       "m4_sweeps_small": it, "m4_misfit_small": float(stats[0])}
        log("CPU %s: asm %.0f s, solve %.0f s (%d sweeps on %d elements, scaled by (n/%d)^2)"
            % (label, t_asm, t_solve, it, ns, ns))
    OR.set_threads(threads)
    st = legs["single_thread"]
    return {"value": st["value"], "unit": "Q-blocks/s (end to end: assemble + Solve prec %.0e)" % args.prec_large,
            "cores": 1, "kind": "port",
            "sample": ("FP64 oracle (oracle/radora.c) on the large yoke: assembly of %d of %d rows "
                       "extrapolated by row count; method-4 Gauss-Seidel (FP32 dot products, as the "
                       "reference) of a %d-tetrahedron yoke of the same construction to prec %.0e, "
                       "%d sweeps, sweep cost scaled by (%d/%d)^2 for the same sweep count; est. asm "
                       "%.0f s, solve %.0f s single-threaded (the reference's assembly and method 4 "
                       "are serial)" % (rows, n, ns, args.prec_large, st["m4_sweeps_small"], n, ns,
                                         st["asm_s"], st["solve_s"])),
            "openmp": legs["openmp"]}


# ---------------------------------------------------------------- Halbach ring
def run_halbach(ctx, args, cpu):
    from radiacuda_amd import rad
    from radiacuda_amd import workloads as W
    N = ctx.N
    log("== Halbach ring: building 20,000 hexahedra and 2M points")
    t0 = time.perf_counter()
    rad.UtiDelAll()
    top, pts = W.halbach()
    k = rad.RlxPre(top)
    r = rad.RlxAuto(k, 1e-8, 1000)
    it = rad._reg[k]
    items, P, model = it.items, it.packed, it.model
    M = np.stack([o.M for (o, _k, _s) in items])
    n, npts = len(items), len(pts)
    log("Halbach: relaxed (misfit %.1e, %d it), set up in %.1f s" % (r[0], r[3], time.perf_counter() - t0))
    d_M, d_pts, d_B = ctx.t64(M), ctx.t64(pts), ctx.zeros(npts, 3)
    out = {}
    for label, prec in (("fp64", N.RH_FP64), ("fp32", N.RH_MIXED)):
        def step(phase):
            N.check(ctx.L.rh_field_dev(model.h, _vp(d_M), _vp(d_pts), npts, _vp(d_B), None, prec))

        wall, _phase, kt = timed_steps(ctx, step, args.steps, args.warmup, "Halbach " + label)
        out[label] = dict(ms=1e3 * wall / args.steps, kernel_ms=kt["field"][0], B=d_B.cpu().numpy())
        log("Halbach %s: %.1f ms per 2M-point sweep" % (label, out[label]["ms"]))
    ne, nf, _ni = model.mesh_counts()
    sec = {"elements": n, "points": npts, "mesh_unique_edges": ne, "mesh_unique_faces": nf}
    for label, o in out.items():
        sec[label] = {"fld_points_per_s": npts / (o["ms"] * 1e-3), "sweep_ms": o["ms"],
                      "field_kernel_ms": o["kernel_ms"]}
    if "fp64" in out and "fp32" in out:
        B64 = out["fp64"]["B"]
        sec["fp32"]["B_err_vs_fp64"] = float(np.abs(out["fp32"]["B"] - B64).max() / np.abs(B64).max())
        sec["fp32"]["speedup_vs_fp64"] = out["fp64"]["ms"] / out["fp32"]["ms"]
    if cpu:
        from oracle import oracle as OR
        OR.set_randomization(False)
        om = OR.prepare_model(P)
        samp = np.sort(np.random.default_rng(3).choice(npts, args.cpu_points, replace=False))
        OR.set_threads(1)
        tc = time.perf_counter()
        Bo, _ = OR.field(om, M.ravel(), pts[samp])
        t1 = (time.perf_counter() - tc) / len(samp)
        OR.set_threads(len(os.sched_getaffinity(0)))
        sec["fp64"]["B_err_vs_oracle_sampled"] = float(np.abs(out["fp64"]["B"][samp] - Bo).max()
                                                       / np.abs(Bo).max())
        sec["fp32"]["B_err_vs_oracle_sampled"] = float(np.abs(out["fp32"]["B"][samp] - Bo).max()
                                                       / np.abs(Bo).max())
        sec["cpu_single_thread_points_per_s"] = 1.0 / t1
        sec["cpu_sample"] = "oracle field at %d of %d points, one thread" % (len(samp), npts)
    rad.UtiDelAll()
    return sec


# ---------------------------------------------------------------------- C-yoke
def run_c_yoke(ctx, args, cpu):
    from radiacuda_amd import _abi
    from radiacuda_amd import objects as O
    from radiacuda_amd import rad
    from radiacuda_amd import workloads as W
    N = ctx.N
    log("== C-yoke: building 8,000 tetrahedra and the 1M-point grid")
    rad.UtiDelAll()
    top, iron, coil, info = W.c_yoke()
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P, stream=ctx.sh)
    X = model.colloc()
    coils = Coils(ctx, O.pack_currents(list(O.walk(rad._reg, coil))))
    pts = W.gap_grid(info, args.grid)
    npts = len(pts)
    d_pts, d_X, d_mats = ctx.t64(pts), ctx.t64(X), ctx.t64(mats)
    d_hext, d_M, d_H, d_B = ctx.zeros(n, 3), ctx.zeros(n, 3), ctx.zeros(n, 3), ctx.zeros(npts, 3)
    out = {}
    for label, prec in (("fp64", N.RH_FP64), ("mixed", N.RH_MIXED)):
        if args.only and label != args.only:
            continue
        Q = model.assemble(prec, dense_rows=args.dense_rows)
        planes = Q.format()[0]
        opts = _abi.RelaxOpts(args.prec, args.max_iter, 0, 0.0, 9, 0)
        res = _abi.RelaxResult()

        def step(phase):
            t = time.perf_counter()
            Q.reassemble()
            t1 = time.perf_counter()
            d_hext.zero_()
            coils.add(d_X, n, d_hext)
            N.check(ctx.L.rh_relax_dev(Q.h, _vp(d_mats), _vp(d_hext), _vp(d_M), _vp(d_H),
                                       ctypes.byref(opts), ctypes.byref(res)))
            t2 = time.perf_counter()
            N.check(ctx.L.rh_field_dev(model.h, _vp(d_M), _vp(d_pts), npts, _vp(d_B), None, prec))
            coils.add(d_pts, npts, d_B)
            ctx.sync()
            t3 = time.perf_counter()
            _acc(phase, "asm", t1 - t)
            _acc(phase, "solve", t2 - t1)
            _acc(phase, "fld", t3 - t2)

        wall, phase, kt = timed_steps(ctx, step, args.steps, args.warmup, "C-yoke " + label)
        K = args.steps
        out[label] = dict(ms_per_step=1e3 * wall / K, asm_ms=1e3 * phase["asm"] / K,
                          solve_s=phase["solve"] / K, fld_ms=1e3 * phase["fld"] / K, iters=res.iters,
                          misfit=res.misfit, refined=res.refined,
                          kernel_ms={f: kt[f][0] for f in kt}, B=d_B.cpu().numpy(), M=d_M.cpu().numpy(),
                          planes=planes)
        Q.close()
        r = out[label]
        log("C-yoke %s: %.1f ms/step (asm %.2f ms, solve %.3f s in %d it, fld %.1f ms)"
            % (label, r["ms_per_step"], r["asm_ms"], r["solve_s"], r["iters"], r["fld_ms"]))
    ne, nf, _ni = model.mesh_counts()
    field_form = model.field_form()
    model.close()
    sec = {"elements": n, "points": npts, "prec": args.prec, "mesh_unique_edges": ne,
           "mesh_unique_triangles": nf}
    for label, r in out.items():
        fl = r["kernel_ms"]["field"]
        flops = float(npts) * (ne * FLOP_MESH_EDGE + nf * FLOP_MESH_TRI) if field_form == 2 else \
            float(npts) * n * FLOP_TET_FIELD
        sec[label] = {"end_to_end_ms": r["ms_per_step"], "end_to_end_points_per_s": npts / (r["ms_per_step"] * 1e-3),
                      "asm_qblocks_per_s": n * n / (r["asm_ms"] * 1e-3), "asm_kernel_ms": r["kernel_ms"]["asm"],
                      "solve_s": r["solve_s"], "solve_iters": r["iters"], "solve_misfit": r["misfit"],
                      "matvec_kernel_ms": r["kernel_ms"]["matvec"], "matrix_planes": r["planes"],
                      "fld_points_per_s": npts / (r["fld_ms"] * 1e-3), "fld_kernel_ms": fl,
                      "fld_declared_TFLOPs": flops / (fl * 1e-3) / 1e12 if fl > 0 else None}
    if "fp64" in out:
        f = sec["fp64"]["fld_declared_TFLOPs"]
        sec["fp64"]["fld_valu_frac_declared"] = f / PEAK_FP64_VALU if f else None
    if "fp64" in out and "mixed" in out:
        r64, rmx = out["fp64"], out["mixed"]
        sec["mixed"]["speedup_vs_fp64"] = r64["ms_per_step"] / rmx["ms_per_step"]
        sec["mixed"]["B_err_vs_fp64"] = float(np.abs(rmx["B"] - r64["B"]).max() / np.abs(r64["B"]).max())
        sec["mixed"]["refinement_passes"] = rmx["refined"]
    return sec


def _oracle_currents(OR, cur, pts):
    """B of packed current sources by the oracle (lab frame)."""
    B = np.zeros((len(pts), 3))
    for k in range(len(cur["kind"])):
        p = cur["par"][k]
        M, V, s = cur["cM"][k], cur["cV"][k], cur["cS"][k]
        for i, x in enumerate(pts):
            xl = np.linalg.solve(M, x - V)
            if cur["kind"][k] == 0:
                b, _ = OR.reccur_BH(xl, p[0:3], p[3:6], p[6:9])
            else:
                b, _ = OR.arccur_BH(xl, p[0:3], p[3], p[4], p[5], p[6], p[7], p[8], int(p[9]))
            B[i] += s * (M @ b)
    return B


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("all", "large_yoke", "halbach", "c_yoke"), default="all",
                    help="one workload only (profiler runs); 'all' is the benchmark")
    ap.add_argument("--grid", type=int, default=100, help="C-yoke field grid points per axis")
    ap.add_argument("--prec", type=float, default=1e-4, help="C-yoke relaxation precision (T)")
    ap.add_argument("--prec-large", type=float, default=1e-5, help="large-yoke relaxation precision (T)")
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline legs")
    ap.add_argument("--cpu-rows", type=int, default=16, help="large-yoke rows in the CPU sample")
    ap.add_argument("--cpu-points", type=int, default=16, help="Halbach points in the CPU sample")
    ap.add_argument("--dense-rows", action="store_true",
                    help="store row-major 3x3 blocks even where the symmetric-block layout applies")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="time one precision only (profiler runs); prints no JSON line")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the steps, copy a MIB-sized device buffer twice (a kernel of "
                         "known HBM traffic, for calibrating profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the multi-GPU sharding of the matrix rows is a "
              "design, not built; run with --gpus 1", file=sys.stderr)
        return 2
    ctx = Ctx()
    cpu = not args.no_cpu and not args.only and args.workload == "all"
    line = {"metric": METRIC}
    large = halb = cyoke = None
    r_large = {}
    if args.workload in ("all", "large_yoke"):
        large, r_large, P_large, mats_large = run_large_yoke(ctx, args, cpu)
    if args.workload in ("all", "halbach") and not args.only:
        halb = run_halbach(ctx, args, cpu)
    if args.workload in ("all", "c_yoke"):
        cyoke = run_c_yoke(ctx, args, cpu)
    if args.calib_copy > 0:
        nel = args.calib_copy * (1 << 20) // 8
        src = ctx.torch.ones(nel, dtype=ctx.torch.float64, device=ctx.dev)
        dst = ctx.torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        ctx.sync()
        log("calibration copy: %d MiB read + %d MiB written, twice" % (args.calib_copy, args.calib_copy))
    if args.only or args.workload != "all":
        return 0
    n = large["elements"]
    r64 = r_large["fp64"]
    ms = r64["ms_per_step"]
    cpu_base = cpu_baseline_large(P_large, mats_large, n, r64, args) if cpu else None
    # roofline of the headline's dominant kernel: the FP64 matrix pass
    n3 = 3 * n
    mv_ms = r64["kernel_ms"]["matvec"]
    bytes_pass = r64["bytes_pass"]
    achieved = bytes_pass / (mv_ms * 1e-3) / 1e9
    kname = ("k_matvec_sym_f64<%d> (FP64 matrix pass, large yoke)" % r64["planes"] if r64["planes"]
             else "k_matvec_f64<4> (FP64 matrix pass, large yoke)")
    traffic = None
    traffic_src = None
    try:
        with open(os.path.join(ROOT, "profiles", "matvec_traffic.json")) as fh:
            tr = json.load(fh)
        if tr.get("workload_elements") == n and tr.get("planes") == r64["planes"]:
            traffic = float(tr["traffic_bytes"])
            traffic_src = "profiles/matvec_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)"
    except (OSError, ValueError, KeyError):
        traffic = None
    line.update({
        "value": n * n / (ms * 1e-3),
        "unit": "Q-blocks/s (large yoke, FP64: assemble + Solve to prec %.0e, end to end)" % args.prec_large,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded meshes and point sets, no files)",
        "config": {"workload": "large yoke: C-shaped dipole yoke of %d jittered tetrahedra (5 per hex "
                               "cell, +-10%% interior jitter), MatSatIsoFrm steel, two racetrack coils; "
                               "assemble the %dx%d matrix + Solve(prec %.0e); also timed: Halbach ring "
                               "field (20,000 hexahedra, 2M points) and the 8,000-tetrahedron C-yoke end "
                               "to end with a 1M-point map" % (n, n3, n3, args.prec_large),
                   "elements": n, "matrix": "%dx%d" % (n3, n3)},
        "large_yoke": large, "halbach": halb, "c_yoke": cyoke,
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": achieved, "peak": PEAK_HBM, "unit": "GB/s", "frac": achieved / PEAK_HBM,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_launch": bytes_pass, "avg_launch_ms": mv_ms},
        "cpu_baseline": cpu_base,
    })
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
