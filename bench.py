#!/usr/bin/env python3
"""Benchmark of the MI355X MG-PCG hot path.

Workload (one "step" = one full solve to relative residual 1e-8 from x = 0, inputs resident in
HBM): 3D Poisson, Q1 hexahedra on the unit cube, level 7 (128^3 cells, 2,146,689 DOF),
f = 3 pi^2 prod sin(pi x_i), zero Dirichlet; PCG (FP64 outer) preconditioned by an MG V(4,4)
cycle with damped Jacobi (omega 0.5) and an exact 2x2x2-cell coarse solve.  The timed solver
runs its V-cycle in FP32 inside the FP64 PCG (ConvertPrecond semantics); the FP64 path is timed
beside it.  Synthetic data: the system is assembled by the package's own setup.

Prints ONE JSON line (rank 0).  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--inner", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--fmt", default="csr", choices=["csr", "sell"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from feat3_amd import problems as pb
    from feat3_amd.solver import MGPCG

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")

    t = time.time()
    H = pb.hierarchy(3, 1, args.level, 1)
    b = pb.filtered_rhs(H, 0)
    n = H.n
    log(f"setup: {len(H.levels)} levels, n={n}, nnz={H.levels[0].A.nnz} ({time.time() - t:.1f} s)")

    db = torch.from_numpy(b).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)

    def run(solver, steps, warmup):
        for _ in range(warmup):
            x.zero_()
            solver.correct(x, db)
        torch.cuda.synchronize()
        times, iters = [], []
        for _ in range(steps):
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, it = solver.correct(x, db)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            iters.append(it)
            assert st == 2, f"solve did not converge (status {st})"
        return times, iters

    t = time.time()
    s = MGPCG(H, inner=args.inner, fmt=args.fmt).init()
    log(f"init {args.inner}/{args.fmt}: {time.time() - t:.1f} s, info {s.info()}")
    times, iters = run(s, args.steps, args.warmup)
    total = sum(times)
    ms_per_step = 1e3 * total / args.steps
    it_mean = float(np.mean(iters))
    value = n * sum(iters) / total
    hist = s.history()
    log(f"{args.inner}: {ms_per_step:.3f} ms/solve, iterations {iters}, final rel. residual {hist[-1] / hist[0]:.3e}")
    xm = x.cpu().numpy().copy()

    # dominant kernel: the finest-level damped-Jacobi sweep, timed on the solver's own stream
    sweep_ms, sweep_bytes = s.time_kernel(0, 50)
    defect_ms, defect_bytes = s.time_kernel(1, 50)
    vc_ms, _ = s.time_kernel(3, 10)
    info = s.info()
    achieved = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    log(f"sweep {sweep_ms * 1e3:.1f} us, {sweep_bytes / 1e6:.1f} MB -> {achieved:.0f} GB/s; "
        f"defect {defect_ms * 1e3:.1f} us -> {defect_bytes / defect_ms / 1e6:.0f} GB/s; V-cycle {vc_ms:.3f} ms")

    # FP64 path beside it
    fp64 = None
    if not args.no_fp64 and args.inner != "f64":
        s64 = MGPCG(H, inner="f64", fmt=args.fmt).init()
        t64, i64 = run(s64, max(1, args.steps // 2), 1)
        x64 = x.cpu().numpy().copy()
        sw64_ms, sw64_bytes = s64.time_kernel(0, 30)
        fp64 = dict(ms_per_solve=1e3 * sum(t64) / len(t64), iterations=i64,
                    dof_iter_per_s=n * sum(i64) / sum(t64),
                    sweep_GBs=sw64_bytes / (sw64_ms * 1e-3) / 1e9,
                    rel_diff_mixed_vs_fp64=float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)))
        s64.done()
        log(f"fp64: {fp64}")
    s.done()

    # CPU baseline: the oracle's FP64 PCG-MG on one host core, one full solve of the same
    # system (bounded sample: one solve, ~10-30 s)
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        S = po.System.from_hierarchy(H)
        xo = np.zeros(n)
        t0 = time.perf_counter()
        st, it, _ = S.solve(xo, b, po.params())
        dt = time.perf_counter() - t0
        cpu = {"value": n * it / dt, "unit": "DOF*iter/s", "cores": 1, "kind": "port",
               "sample": f"one full FP64 PCG-MG solve of the same system ({it} iterations, {dt:.1f} s, "
                         "single-threaded oracle restating FEAT's generic CPU loops)"}
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "time to rel. residual 1e-8 & DOF·iter/s, FP32-inner vs FP64; SpMV GB/s",
        "value": value,
        "unit": "DOF*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.inner == "f32" else ("f64" if args.inner == "f64" else "f16-values/f32"),
        "data": "synthetic (assembled 3D Q1 Poisson system, sine-bubble load)",
        "config": {"workload": f"3D Poisson Q1 L{args.level}, {n} DOF, PCG(FP64)+MG V(4,4) Jacobi 0.5, "
                               f"{args.inner} V-cycle, {args.fmt}",
                   "dofs": n, "levels": len(H.levels), "tol_rel": 1e-8},
        "time_to_tol_ms": ms_per_step,
        "iterations": it_mean,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sweep (finest level)", "bytes_per_launch": sweep_bytes,
                     "avg_launch_us": sweep_ms * 1e3},
        "vcycle_ms": vc_ms,
        "vcycle_kernels": info["kernels_per_vcycle"],
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["ms_per_solve"] / ms_per_step) if fp64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
