#!/usr/bin/env python3
"""Benchmark of the three paths on one MI355X: interaction-matrix assembly,
relaxation solve and field evaluation, end to end, on the C-shaped dipole yoke
(8,000 jittered tetrahedra, nonlinear steel, two racetrack coils, B on a
100^3 grid in the gap).

A step = assemble the 24,000 x 24,000 interaction matrix + relax from the
remanent state + evaluate B (magnetized elements and coils) at 1,000,000
points, with every input already resident in HBM.  Two precisions are timed:
FP64 (FP64 matrix storage, FP64 integrals) and mixed (FP32 matrix storage,
FP32 integrals of the field evaluation, FP64 accumulation and solve).

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
# Declared algorithmic FLOPs per observation point, counting + - * / sqrt as 1,
# log as 20 and atan2 as 30, from the integrals as written:
# - per-element form (radiacuda_amd/csrc/rh_math.hpp), per (point, tetrahedron)
#   pair: 4 sqrt, 6 log, 6 div, 4 atan2;
# - mesh form (radiacuda_amd/csrc/rh_mesh.hpp), per (point, unique edge):
#   2 sqrt, 1 div, 1 log; per (point, unique triangle): 3 sqrt, 1 atan2.
FLOP_TET_FIELD = 821
FLOP_TET_BLOCK = 794
FLOP_MESH_EDGE = 74
FLOP_MESH_TRI = 93
PEAK_FP64_VALU = 78.6   # TFLOP/s, MI355X vector FP64
PEAK_FP32_VALU = 157.3  # TFLOP/s, MI355X vector FP32
PEAK_HBM = 8000.0       # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=100, help="field grid points per axis")
    ap.add_argument("--prec", type=float, default=1e-4, help="relaxation precision (T)")
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=48, help="rows/points in the CPU sample")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="time one precision only (profiler runs); prints no JSON line")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the steps, copy a MIB-sized device buffer twice (a kernel of "
                         "known HBM traffic, for calibrating profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every box has a single MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    from radiacuda_amd import _abi
    from radiacuda_amd import _native as N
    from radiacuda_amd import objects as O
    from radiacuda_amd import rad
    from radiacuda_amd import workloads as W

    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    L = N.lib()

    log("building C-yoke model")
    t0 = time.perf_counter()
    top, iron, coil, info = W.c_yoke()
    items = list(O.walk(rad._reg, iron))
    P = O.pack_elements(items)
    n = len(items)
    mats = np.stack([o.mat.record(o.M) for (o, _k, _s) in items])
    model = N.Model(P, stream=sh)
    X = model.colloc()
    cur = O.pack_currents(list(O.walk(rad._reg, coil)))
    pts = W.gap_grid(info, args.grid)
    npts = len(pts)
    assert all(np.allclose(P["cM"][c], np.eye(3)) for c in range(len(P["cM"])))
    log("model: %d tetrahedra, %d points, built in %.1f s" % (n, npts, time.perf_counter() - t0))

    f64 = dict(dtype=torch.float64, device=dev)
    d_pts = torch.tensor(pts, **f64)
    d_X = torch.tensor(X, **f64)
    d_mats = torch.tensor(mats, **f64)
    d_hext = torch.zeros((n, 3), **f64)
    d_M = torch.zeros((n, 3), **f64)
    d_H = torch.zeros((n, 3), **f64)
    d_B = torch.zeros((npts, 3), **f64)
    kind = np.ascontiguousarray(cur["kind"], dtype=np.int32)
    par = np.ascontiguousarray(cur["par"])
    tm, tv, ts = (np.ascontiguousarray(cur[k]) for k in ("cM", "cV", "cS"))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    vp = ctypes.c_void_p
    torch.cuda.synchronize()

    def currents(d_points, count, d_out):
        N.check(L.rh_current_field_dev(0, sh, len(kind), kind.ctypes.data_as(ip),
                                       par.ctypes.data_as(dp), tm.ctypes.data_as(dp),
                                       tv.ctypes.data_as(dp), ts.ctypes.data_as(dp),
                                       vp(d_points.data_ptr()), count, vp(d_out.data_ptr())))

    results = {}
    for label, prec in (("fp64", N.RH_FP64), ("mixed", N.RH_MIXED)):
        if args.only and label != args.only:
            continue
        log("== %s: allocating and assembling the matrix" % label)
        Q = model.assemble(prec)
        opts = _abi.RelaxOpts(args.prec, args.max_iter, 0, 0.0, 9, 0)
        res = _abi.RelaxResult()
        phase = {"asm": 0.0, "hext": 0.0, "solve": 0.0, "fld": 0.0}

        def step(timed):
            t = time.perf_counter()
            Q.reassemble()
            t1 = time.perf_counter()
            d_hext.zero_()
            currents(d_X, n, d_hext)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            N.check(L.rh_relax_dev(Q.h, vp(d_mats.data_ptr()), vp(d_hext.data_ptr()),
                                   vp(d_M.data_ptr()), vp(d_H.data_ptr()), ctypes.byref(opts),
                                   ctypes.byref(res)))
            t3 = time.perf_counter()
            N.check(L.rh_field_dev(model.h, vp(d_M.data_ptr()), vp(d_pts.data_ptr()), npts,
                                   vp(d_B.data_ptr()), None, prec))
            currents(d_pts, npts, d_B)
            torch.cuda.synchronize()
            t4 = time.perf_counter()
            if timed:
                phase["asm"] += t1 - t
                phase["hext"] += t2 - t1
                phase["solve"] += t3 - t2
                phase["fld"] += t4 - t3

        for w in range(args.warmup):
            step(False)
            log("warmup %d done (solve %d iterations, misfit %.3e)" % (w, res.iters, res.misfit))
        N.set_timing(True)
        torch.cuda.synchronize()
        tstart = time.perf_counter()
        for s in range(args.steps):
            step(True)
            log("step %d done" % s)
        torch.cuda.synchronize()
        wall = time.perf_counter() - tstart
        kt = {f: N.kernel_time(f) for f in ("asm", "field", "matvec", "refine")}
        N.set_timing(False)
        K = args.steps
        results[label] = dict(
            wall_s=wall, ms_per_step=1e3 * wall / K,
            asm_ms=1e3 * phase["asm"] / K, hext_ms=1e3 * phase["hext"] / K,
            solve_s=phase["solve"] / K, fld_ms=1e3 * phase["fld"] / K,
            iters=res.iters, misfit=res.misfit, converged=bool(res.converged),
            refined=res.refined,
            kernel_ms={f: kt[f][0] for f in kt}, kernel_counts={f: kt[f][1] for f in kt},
            B=d_B.cpu().numpy(), M=d_M.cpu().numpy())
        Q.close()
        r = results[label]
        log("%s: %.1f ms/step (asm %.2f ms, solve %.3f s in %d it, fld %.1f ms)"
            % (label, r["ms_per_step"], r["asm_ms"], r["solve_s"], r["iters"], r["fld_ms"]))

    if args.calib_copy > 0:
        nel = args.calib_copy * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read + %d MiB written, twice" % (args.calib_copy, args.calib_copy))
    if args.only:
        return 0
    r64, rmx = results["fp64"], results["mixed"]
    scale = np.abs(r64["B"]).max()
    err_mixed = float(np.abs(rmx["B"] - r64["B"]).max() / scale)
    err_M = float(np.abs(rmx["M"] - r64["M"]).max() / max(np.abs(r64["M"]).max(), 1e-300))

    # FP64 GPU vs the FP64 CPU oracle on a sample of points (same magnetization)
    cpu = None
    err_oracle = None
    if not args.no_cpu:
        from oracle import oracle as OR
        OR.set_randomization(False)
        threads = len(os.sched_getaffinity(0))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        rng = np.random.default_rng(0)
        samp = rng.choice(npts, args.cpu_sample, replace=False)
        log("CPU leg: oracle on %d of %d points, %d of %d rows, %d threads"
            % (len(samp), npts, args.cpu_sample, n, threads))
        om = OR.prepare_model(P)
        tc = time.perf_counter()
        Bs, _ = OR.field(om, r64["M"].ravel(), pts[samp])
        t_fld = (time.perf_counter() - tc) * npts / len(samp)
        Bs = Bs + _oracle_currents(OR, cur, pts[samp])
        err_oracle = float(np.abs(Bs - r64["B"][samp]).max() / np.abs(Bs).max())
        tc = time.perf_counter()
        Qs = OR.assemble(om, rows=(0, args.cpu_sample))
        t_asm = (time.perf_counter() - tc) * n / args.cpu_sample
        # one Gauss-Seidel sweep costs 9 n^2 FP32 multiply-adds; time a sweep on
        # the sampled rows and charge the GPU solve's iteration count
        sub = Qs[:, :3 * args.cpu_sample].astype(np.float32)
        tc = time.perf_counter()
        OR.relax_m4(sub, mats[:args.cpu_sample], np.zeros((args.cpu_sample, 3)), 0.0, 5)
        t_sweep = (time.perf_counter() - tc) / 5 * (n / args.cpu_sample) ** 2
        t_solve = t_sweep * r64["iters"]
        t_cpu = t_asm + t_solve + t_fld
        cpu = {"value": npts / t_cpu, "unit": "Fld points/s (end-to-end: assemble+solve+Fld)",
               "cores": threads, "kind": "port",
               "sample": ("FP64 oracle (radora.c, OpenMP) extrapolated linearly: field on %d of "
                          "%d points, assembly of %d of %d rows, one Gauss-Seidel sweep on a "
                          "%dx%d sub-block scaled by (n/%d)^2 and by the GPU solve's %d "
                          "iterations; est. asm %.1f s, solve %.1f s, fld %.1f s"
                          % (len(samp), npts, args.cpu_sample, n, args.cpu_sample,
                             args.cpu_sample, args.cpu_sample, r64["iters"], t_asm, t_solve,
                             t_fld))}

    # roofline of the dominant kernel (field evaluation, FP64 integrals)
    fld_ms = r64["kernel_ms"]["field"]
    if model.field_form() == 2:
        ne, nf, _ni = model.mesh_counts()
        flops = float(npts) * (ne * FLOP_MESH_EDGE + nf * FLOP_MESH_TRI)
        fld_kernel = ("k_field_mesh<double> (FP64 integrals over %d unique edges and %d unique "
                      "triangles)" % (ne, nf))
    else:
        flops = float(npts) * n * FLOP_TET_FIELD
        fld_kernel = "k_field_tet<double> (FP64 integrals)"
    achieved = flops / (fld_ms * 1e-3) / 1e12 if fld_ms > 0 else 0.0
    # HBM bytes per launch of the dominant kernel, measured by rocprofv3
    # (FETCH_SIZE / WRITE_SIZE passes, calibrated on a device copy) and kept
    # in profiles/field_traffic.json for this kernel and workload
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "field_traffic.json")) as fh:
            tr = json.load(fh)
        if fld_kernel.startswith(tr["kernel"]) and args.grid == 100 and n == 8000:
            traffic = float(tr["traffic_bytes"])
    except (OSError, ValueError, KeyError):
        traffic = None
    asm_ms = r64["kernel_ms"]["asm"]
    mv_ms = r64["kernel_ms"]["matvec"]
    n3 = 3 * n
    ms = r64["ms_per_step"]
    line = {
        "metric": METRIC,
        "value": npts / (ms * 1e-3),
        "unit": "Fld points/s (end-to-end: assemble+solve+Fld of a 1M-point map)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded C-yoke mesh, no files)",
        "config": {"workload": "C-shaped dipole yoke: %d jittered tetrahedra (5 per hex cell, "
                               "+-10%% interior jitter), MatSatIsoFrm steel, two racetrack coils; "
                               "assemble + Solve(prec %.0e) + B on a %d^3 gap grid"
                               % (n, args.prec, args.grid),
                   "elements": n, "points": npts, "matrix": "%dx%d" % (n3, n3)},
        "fp64": {"asm_qblocks_per_s": n * n / (r64["asm_ms"] * 1e-3),
                 "asm_kernel_ms": asm_ms, "solve_s": r64["solve_s"], "solve_iters": r64["iters"],
                 "solve_misfit": r64["misfit"], "matvec_kernel_ms": mv_ms,
                 "matvec_GBps": 8.0 * n3 * n3 / (mv_ms * 1e-3) / 1e9 if mv_ms > 0 else None,
                 "fld_points_per_s": npts / (r64["fld_ms"] * 1e-3),
                 "fld_kernel_ms": fld_ms, "end_to_end_ms": ms,
                 "err_vs_oracle_sampled": err_oracle},
        "mixed": {"asm_qblocks_per_s": n * n / (rmx["asm_ms"] * 1e-3),
                  "asm_kernel_ms": rmx["kernel_ms"]["asm"], "solve_s": rmx["solve_s"],
                  "solve_iters": rmx["iters"], "solve_misfit": rmx["misfit"],
                  "fp64_refinement_passes": rmx["refined"],
                  "refine_field_ms": rmx["kernel_ms"]["refine"],
                  "matvec_kernel_ms": rmx["kernel_ms"]["matvec"],
                  "matvec_GBps": (4.0 * n3 * n3 / (rmx["kernel_ms"]["matvec"] * 1e-3) / 1e9
                                  if rmx["kernel_ms"]["matvec"] > 0 else None),
                  "fld_points_per_s": npts / (rmx["fld_ms"] * 1e-3),
                  "fld_kernel_ms": rmx["kernel_ms"]["field"], "end_to_end_ms": rmx["ms_per_step"],
                  "speedup_vs_fp64": ms / rmx["ms_per_step"],
                  "B_err_vs_fp64": err_mixed, "M_err_vs_fp64": err_M},
        "roofline": {"bound": "valu", "kernel": fld_kernel,
                     "achieved": achieved, "peak": PEAK_FP64_VALU, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_VALU, "traffic": traffic,
                     "flops_per_launch": flops, "avg_launch_ms": fld_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _oracle_currents(OR, cur, pts):
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


if __name__ == "__main__":
    sys.exit(main())
