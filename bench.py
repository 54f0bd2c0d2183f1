#!/usr/bin/env python3
"""Benchmark of the MI355X MG-PCG hot path.

Workload (one "step" = one full solve to relative residual 1e-8 from x = 0, inputs resident in
HBM, setup excluded): 3D Poisson, Q1 hexahedra on the unit cube, level 7 (128^3 cells,
2,146,689 DOF, 57,066,625 nnz), f = 3 pi^2 prod sin(pi x_i), zero Dirichlet; PCG (FP64 outer:
residual, update, dot products) preconditioned by an MG V(4,4) cycle with Richardson(1.0,
Jacobi 0.5) smoothing and an exact FP64 solve on the 2x2x2-cell coarse grid.  The headline
(value) runs the V-cycle in FP32 inside the FP64 PCG (ConvertPrecond semantics); the FP64 path
and the FP16-stored-values/FP32-arithmetic V-cycle are timed beside it on the same system.
Synthetic data: the system is assembled by the package's own setup.

Prints ONE JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec

# HBM bytes per launch of the finest-level sweep, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (separate runs of this script) and scaled by the same counters on a plain 512 MiB
# device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); profiles/k_sweep_3d_q1_l7_f32.md,
# profiles/solve_kernels_2d_q2_l9.md and profiles/batch_k_sweep_vector32_f32.md hold the commands
# and raw values (the two scalar workloads' bands are value-coded).  Keyed by (workload, level, inner, format); absent -> null.
MEASURED_SWEEP_TRAFFIC = {
    ("poisson3d", 7, "f32", "banded"): (2 * 25674.4 + 8385.5) * 1024,
    ("poisson2d", 9, "f32", "banded"): (2 * 14486.4 + 4104.0) * 1024,
    ("vector32", 6, "f32", "banded"): (2 * 288738.1 + 102984.4) * 1024,
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def band_nnz(A, inner, upper=False):
    """Entries the solver's copy of a level matrix keeps in the given inner precision (every format):
    those with |a_ij| > eps^2 |a_ii| (eps of the inner arithmetic; binary16 storage computes in
    FP32; in FP64 that is every stored nonzero);
    upper: only those with j >= i, which is what the symmetric band storage holds."""
    import numpy as np
    rows = np.repeat(np.arange(A.rows), np.diff(A.row_ptr.astype(np.int64)))
    cols = A.col_ind.astype(np.int64)
    diag = np.zeros(A.rows)
    on = cols == rows
    diag[rows[on]] = A.val[on]
    eps = np.finfo(np.float64 if inner == "f64" else np.float32).eps
    keep = np.abs(A.val) > eps * eps * np.abs(diag[rows])
    if upper:
        keep &= cols >= rows
    return int(np.count_nonzero(keep))


def sweep_algorithmic_bytes(info, n, nnz, nchunks):
    """Fused damped-Jacobi sweep x_out = x + dinv (b - A x): the CSR defect formula over the
    unpadded nnz (the entries the level copy keeps, band_nnz) with the stored value and index
    sizes -- for SELL-C the chunk descriptors (cs, cl) replace row_ptr, the banded format stores no
    indices at all (its offsets, 4 bytes per diagonal, replace them; for the symmetric band storage
    nnz counts the entries on and above the diagonal: the lower half is read from the L2) -- x
    counted once, plus dinv, b and x_out.  The value-coded band reads 4-bit codes instead of the
    values; its bytes are the solver's own count (time_kernel), this formula then gives the
    value-stored equivalent."""
    s_v = info["sizeof_value"]
    s_x = info["sizeof_inner"]
    if info.get("banded"):
        return nnz * s_v + 4 * info["ndiag"] + n * s_x * 4
    if info.get("sell_chunks"):
        s_i = 2 if info.get("index16", 1) else 4
        return nnz * (s_v + s_i) + 8 * nchunks + n * s_x * 4
    return nnz * (s_v + 4) + 4 * (n + 1) + n * s_x * 4


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    times, outs = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = fn()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        outs.append(o)
    return times, outs


def _host_cpu():
    """CPU model and the threads visible to this process (for the cpu_baseline record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}; {len(os.sched_getaffinity(0))} threads visible, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}"


def _common(args, metric_value, ms, dtype, data, config, roof, extra, cpu):
    if cpu is not None:
        cpu.setdefault("host", _host_cpu())
    return {
        "metric": "time to rel. residual 1e-8 & DOF·iter/s, FP32-inner vs FP64; SpMV GB/s",
        "value": metric_value, "unit": "DOF*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": data, "config": config, "roofline": roof, **extra, "cpu_baseline": cpu,
    }


def bench_vector32(args):
    """32 right-hand sides (xorshift64* uniform[-1,1], seed 5) of the 3D Q1 level-6 vector Laplacian,
    3x3-block BCSR (274,625 block rows), solved in one batched PCG-MG launch sequence.  One step
    = one batched solve of all 32 systems to relative residual 1e-8 from x = 0."""
    import numpy as np
    import torch
    from feat3_amd import problems as pb
    from feat3_amd.batch import BatchMGPCG

    K, lvl = 32, args.level if args.level != 7 else 6
    t = time.time()
    H = pb.hierarchy(3, 1, lvl, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    B = pb.random_rhs(H.n, 3, K, H.levels[0].bnd, seed=5)
    nodes, nnzb = H.n, H.levels[0].A.nnz
    log(f"setup vector32: {len(H.levels)} levels, {nodes} block rows, {nnzb} blocks ({time.time() - t:.1f} s)")
    dB = torch.from_numpy(B).cuda()
    X = torch.zeros_like(dB)
    res = {}
    for inner in ([args.inner] + ([] if args.no_variants else [v for v in ("f64", "f32") if v != args.inner])):
        s = BatchMGPCG(H, K, values=vals, block=3, inner=inner).init()

        def step():
            X.zero_()
            st, it = s.correct(X, dB)
            assert all(int(v) == 2 for v in st), st
            return it
        times, its = _timed(step, args.steps if inner == args.inner else max(2, args.steps // 2), args.warmup)
        sweep_ms, sweep_b = s.time_kernel(0, 20)
        vc_ms, _ = s.time_kernel(3, 5)
        it = its[-1]
        r = dict(ms_per_solve=1e3 * float(np.mean(times)), iterations=[int(v) for v in it],
                 dof_iter_per_s=3 * nodes * float(np.sum(it)) / float(np.mean(times)),
                 sweep_us=sweep_ms * 1e3, sweep_bytes=sweep_b, sweep_GBs=sweep_b / (sweep_ms * 1e-3) / 1e9,
                 vcycle_ms=vc_ms, solution=X.cpu().numpy().copy())
        log(f"vector32 {inner}: {r['ms_per_solve']:.2f} ms/solve, it {min(it)}..{max(it)}, sweep {r['sweep_us']:.1f} us "
            f"{r['sweep_GBs']:.0f} GB/s, V-cycle {vc_ms:.3f} ms")
        res[inner] = r
        s.done()
    head = res[args.inner]
    if "f64" in res and args.inner != "f64":
        x64 = res["f64"]["solution"]
        head["rel_diff_vs_fp64_solution"] = float(np.linalg.norm(head["solution"] - x64) / np.linalg.norm(x64))
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        S = po.System.from_hierarchy(H, block=3, values=vals)
        xo = np.zeros(3 * nodes)
        st, it, _ = S.solve(xo, B[:, :, 0].reshape(-1).copy(), po.params())
        dt = po.lib().orc_last_solve_seconds()
        cpu = {"value": 3 * nodes * it / dt, "unit": "DOF*iter/s", "cores": 1, "kind": "port",
               "sample": f"one of the 32 right-hand sides solved by the oracle's FP64 blocked PCG-MG on one core "
                         f"({it} iterations, {dt:.2f} s solve phase); the reference solves the 32 one after another",
               "time_to_tol_ms_32_rhs": 32 * dt * 1e3}
        if "f64" in res:
            x64 = res["f64"]["solution"][:, :, 0].reshape(-1)
            cpu["rel_diff_gpu_fp64_vs_cpu_rhs0"] = float(np.linalg.norm(x64 - xo) / np.linalg.norm(xo))
    for r in res.values():
        r.pop("solution")
    roof = {"bound": "hbm", "achieved": head["sweep_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": head["sweep_GBs"] / HBM_PEAK_GBS,
            "traffic": MEASURED_SWEEP_TRAFFIC.get(("vector32", lvl, args.inner, "banded")),
            "kernel": "batch k_sweep<B=3>, finest level (fused damped-Jacobi step, 32 systems)",
            "bytes_per_launch": head["sweep_bytes"], "avg_launch_us": head["sweep_us"]}
    cfg = {"workload": f"3D vector Laplacian Q1 level {lvl}, 3x3 BCSR ({nodes} block rows, {nnzb} blocks), "
                       f"32 right-hand sides in one batched PCG(FP64) + MG V(4,4), {args.inner} V-cycle, rel. tol 1e-8",
           "block_rows": nodes, "blocks": nnzb, "rhs": K}
    return _common(args, head["dof_iter_per_s"], head["ms_per_solve"], args.inner,
                   "synthetic (assembled vector Laplacian, xorshift64* right-hand sides)", cfg, roof,
                   {"headline": head, "variants": {k: v for k, v in res.items() if k != args.inner}}, cpu)


def bench_coeff16(args):
    """16 coefficient fields k = exp(g), g Gaussian random fields (sigma 1, squared-exponential,
    length 0.1, seeds 0..15), on the 2D Q1 level-10 mesh; u_obs from the field of seed 1000; loss
    1/2 sum_j |u_j - u_obs|^2.  One step = one differentiable solve: forward (device assembly of all
    levels from k, batched solve) and backward (the same assembly again, adjoint batched solve,
    dL/dk kernel), through torch.autograd.  The persistent solver object is built once beforehand
    (reported separately, not timed)."""
    import numpy as np
    import torch
    from feat3_amd import problems as pb
    from feat3_amd.autograd import DiffusionProblem, cell_gradient, diffusion_solve

    K, lvl = 16, args.level if args.level != 7 else 10
    prob = DiffusionProblem(dim=2, level=lvl, inner=args.inner)
    t = time.time()
    kf = np.exp(np.stack([pb.gaussian_random_field(2, lvl, seed=s) for s in range(K)]))
    kobs = np.exp(pb.gaussian_random_field(2, lvl, seed=1000))
    t_grf = time.time() - t
    kt = torch.from_numpy(kf).cuda()
    t = time.time()
    with torch.no_grad():
        uobs = diffusion_solve(torch.from_numpy(kobs[None, :].copy()).cuda(), prob)[:, 0].clone()
        prob._solver = None          # the 16-field solver below
        diffusion_solve(kt, prob)    # builds the persistent 16-field solver
    torch.cuda.synchronize()
    t_setup = time.time() - t
    log(f"setup coeff16: GRF {t_grf:.1f} s, solver objects {t_setup:.1f} s, n={prob.nodes}")

    def step():
        k = kt.clone().requires_grad_(True)
        u = diffusion_solve(k, prob)
        it1 = prob.last_iters.copy()
        loss = 0.5 * ((u - uobs[:, None]) ** 2).sum()
        loss.backward()
        it2 = prob.last_iters.copy()
        return it1, it2, k.grad, u.detach()
    times, outs = _timed(step, args.steps, args.warmup)
    it1, it2, g, U = outs[-1]
    iters = float(np.sum(it1) + np.sum(it2))
    ms = 1e3 * float(np.mean(times))
    s = prob._solver
    sweep_ms, sweep_b = s.time_kernel(0, 20)
    Ug, gg = U.cpu().numpy(), g.cpu().numpy()
    uo = uobs.cpu().numpy()
    ident = max(abs(kf[j] @ gg[j] + np.dot(Ug[:, j] - uo, Ug[:, j])) / abs(np.dot(Ug[:, j] - uo, Ug[:, j]))
                for j in range(K))
    # the parts: device assembly of all levels + value refresh, and the gradient kernel alone
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(5):
        prob.solver_for(kt)
    torch.cuda.synchronize()
    asm_ms = 1e3 * (time.perf_counter() - t) / 5
    Lam = U.clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        cell_gradient(prob, U, Lam)
    e1.record()
    torch.cuda.synchronize()
    grad_ms = e0.elapsed_time(e1) / 20
    grad_bytes = 2 * prob.nodes * K * 8 + prob.cells * K * 8
    head = dict(ms_per_step=ms, forward_iterations=[int(v) for v in it1], adjoint_iterations=[int(v) for v in it2],
                dof_iter_per_s=prob.nodes * iters / (ms * 1e-3), sweep_us=sweep_ms * 1e3, sweep_bytes=sweep_b,
                sweep_GBs=sweep_b / (sweep_ms * 1e-3) / 1e9, grad_kernel_us=grad_ms * 1e3,
                grad_kernel_GBs=grad_bytes / (grad_ms * 1e-3) / 1e9, assembly_and_refresh_ms=asm_ms,
                max_rel_identity_error=ident, solver_setup_s=t_setup, grf_s=t_grf)
    log(f"coeff16: {ms:.2f} ms/step, fwd it {min(it1)}..{max(it1)}, adj it {min(it2)}..{max(it2)}, "
        f"sweep {head['sweep_us']:.1f} us {head['sweep_GBs']:.0f} GB/s, assembly+refresh {asm_ms:.2f} ms, "
        f"dL/dk {grad_ms * 1e3:.1f} us, identity {ident:.1e}")
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        H0 = pb.hierarchy(2, 1, lvl, 1, kcell=kf[0])
        S = po.System.from_hierarchy(H0)
        xo = np.zeros(prob.nodes)
        st, it, _ = S.solve(xo, prob.b.copy(), po.params())
        dt = po.lib().orc_last_solve_seconds()
        cpu = {"value": prob.nodes * it / dt, "unit": "DOF*iter/s", "cores": 1, "kind": "port",
               "sample": f"the forward solve of field 0 by the oracle's FP64 PCG-MG on one core ({it} iterations, "
                         f"{dt:.2f} s solve phase)"}
    roof = {"bound": "hbm", "achieved": head["sweep_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": head["sweep_GBs"] / HBM_PEAK_GBS, "traffic": None,
            "kernel": "batch k_sweep<B=1, per-system values>, finest level (16 fields)",
            "bytes_per_launch": head["sweep_bytes"], "avg_launch_us": head["sweep_us"]}
    cfg = {"workload": f"2D variable-coefficient diffusion Q1 level {lvl} ({prob.nodes} DOF), 16 coefficient fields: "
                       f"differentiable solve (device assembly, forward + adjoint batched PCG(FP64) + MG V(4,4), "
                       f"{args.inner} V-cycle, rel. tol 1e-8, dL/dk kernel)",
           "dofs": prob.nodes, "fields": K}
    return _common(args, head["dof_iter_per_s"], ms, args.inner,
                   "synthetic (seeded Gaussian random fields, f = 1)", cfg, roof, {"headline": head}, cpu)


def bench_spmv(args):
    """The drop-in Arch::Apply kernels (lafem_hip_apply_*: r = A x, beta = 0) and BLAS-1 at full
    size: the scalar formats on the 3D Q1 level-7 Laplacian (2,146,689 rows, 57,066,625 nnz), the
    blocked ones on the 3D Q1 level-6 vector Laplacian (274,625 block rows of 3x3 blocks).  Each
    kernel: 20 launches between two HIP events on the launching stream after 3 warm-ups; bytes are
    the algorithmic ones (unpadded nnz; x and r once), so padding and re-reads show as lost
    efficiency.  One step = one pass over the whole table."""
    import ctypes
    import numpy as np
    import torch
    from feat3_amd import capi
    from feat3_amd import problems as pb

    capi.call("feat3_hip_initialise", 0)   # MemoryPool<Mem::HIP>::initialise (reduction workspace)

    lvl = args.level
    t = time.time()
    A, _ = pb.laplace(3, 1, lvl)
    Av, _ = pb.laplace(3, 1, max(lvl - 1, 1))
    log(f"setup spmv: n={A.rows}, nnz={A.nnz}; vector Laplacian {Av.rows} block rows ({time.time() - t:.1f} s)")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    keep = []

    def dev(a):
        d = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append(d)
        return ctypes.c_void_p(d.data_ptr())

    def timed(fn, reps=20):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3   # seconds per launch

    n, nnz = A.rows, A.nnz
    rng = np.random.default_rng(0)
    rows = []

    def add(name, fmt, dtype, bytes_, fn, ref_bytes=None):
        sec = timed(fn)
        gbs = bytes_ / sec / 1e9
        row = dict(kernel=name, format=fmt, dtype=dtype, us=sec * 1e6, bytes=int(bytes_), GBs=gbs,
                   frac_of_peak=gbs / HBM_PEAK_GBS)
        if ref_bytes is not None:
            # the reference's own SpMV benchmark count, used_elements * (s_v + s_i) + n * s_v
            # (benchmarks/product_matvec-bench.cpp:130-133), for comparison with its published rates
            row.update(ref_formula_bytes=int(ref_bytes), ref_formula_GBs=ref_bytes / sec / 1e9)
        rows.append(row)
        log(f"{name:38s} {sec * 1e6:9.1f} us {bytes_ / 1e6:9.1f} MB {gbs:7.0f} GB/s  {gbs / HBM_PEAK_GBS:.3f} of peak")

    for dt, npdt, s_v, tag in ((np.float64, "f64", 8, "f64"), (np.float32, "f32", 4, "f32")):
        x = dev(rng.uniform(-1, 1, A.cols).astype(dt))
        r = dev(np.zeros(n, dt))
        val = dev(A.val.astype(dt))
        col, rp = dev(A.col_ind), dev(A.row_ptr)
        add(f"lafem_hip_apply_csr_{tag}_u32", "csr", tag, nnz * (s_v + 4) + (n + 1) * 4 + 2 * n * s_v,
            lambda: capi.call(f"lafem_hip_apply_csr_{tag}_u32", sp, r, 1.0, x, 0.0, r, val, col, rp, n, A.cols, nnz, 0),
            ref_bytes=nnz * (s_v + 4) + n * s_v)
        if tag == "f32":
            hv = dev(A.val.astype(np.float16))
            add("lafem_hip_apply_csr_h16f32_u32", "csr", "f16 values / f32", nnz * (2 + 4) + (n + 1) * 4 + 2 * n * 4,
                lambda: capi.call("lafem_hip_apply_csr_h16f32_u32", sp, r, 1.0, x, 0.0, r, hv, col, rp, n, A.cols, nnz))
        for C, sig in ((32, 1), (64, 64)):
            cs, cl, perm, ecol, evals, rl = pb.to_sell(A, C, sig)
            dcs, dcl, dcol, dval = dev(cs), dev(cl), dev(ecol), dev(evals.astype(dt))
            nch = cl.size
            if sig == 1:
                add(f"lafem_hip_apply_ell_{tag}_u32 (C={C})", "ell", tag, nnz * (s_v + 4) + 8 * nch + 2 * n * s_v,
                    lambda: capi.call(f"lafem_hip_apply_ell_{tag}_u32", sp, r, 1.0, x, 0.0, r, dval, dcol, dcs, dcl, C, n),
                    ref_bytes=nnz * (s_v + 4) + n * s_v)
            else:
                drl, dperm = dev(rl), dev(perm[:n])
                add(f"lafem_hip_apply_sell_{tag}_u32 (C={C}, s={sig})", "sell", tag,
                    nnz * (s_v + 4) + 8 * nch + 8 * n + 2 * n * s_v,
                    lambda: capi.call(f"lafem_hip_apply_sell_{tag}_u32", sp, r, 1.0, x, 0.0, r, dval, dcol, dcs, dcl, drl,
                                      dperm, C, n))
        offs, bval = pb.to_banded(A)
        doff, dbv = dev(offs), dev(bval.astype(dt))
        add(f"lafem_hip_apply_banded_{tag}_u32 ({offs.size} diag.)", "banded", tag, nnz * s_v + 4 * offs.size + 2 * n * s_v,
            lambda: capi.call(f"lafem_hip_apply_banded_{tag}_u32", sp, r, 1.0, x, 0.0, r, dbv, doff, offs.size, n, A.cols))
        # blocked: the vector Laplacian (3x3 blocks a_ij I) and the scalar matrix times 3-vectors
        nb, nnzb = Av.rows, Av.nnz
        xb = dev(rng.uniform(-1, 1, 3 * nb).astype(dt))
        rb = dev(np.zeros(3 * nb, dt))
        bv = dev(pb.block_identity_values(Av, 3).reshape(-1).astype(dt))
        bcol, brp = dev(Av.col_ind), dev(Av.row_ptr)
        add(f"lafem_hip_apply_csrb_{tag}_u32 (3x3)", "bcsr", tag, nnzb * (9 * s_v + 4) + (nb + 1) * 4 + 2 * 3 * nb * s_v,
            lambda: capi.call(f"lafem_hip_apply_csrb_{tag}_u32", sp, rb, 1.0, xb, 0.0, rb, bv, bcol, brp, nb, nb, nnzb, 3, 3))
        sv = dev(Av.val.astype(dt))
        add(f"lafem_hip_apply_csrsb_{tag}_u32 (3)", "csrsb", tag, nnzb * (s_v + 4) + (nb + 1) * 4 + 2 * 3 * nb * s_v,
            lambda: capi.call(f"lafem_hip_apply_csrsb_{tag}_u32", sp, rb, 1.0, xb, 0.0, rb, sv, bcol, brp, nb, nb, nnzb, 3))
        y = dev(rng.uniform(-1, 1, n).astype(dt))
        add(f"lafem_hip_axpy_{tag}", "blas1", tag, 3 * n * s_v,
            lambda: capi.call(f"lafem_hip_axpy_{tag}", sp, r, 2.0, x, y, n))
        res = dev(np.zeros(1, dt))
        add(f"lafem_hip_dot_{tag} (device scalar)", "blas1", tag, 2 * n * s_v,
            lambda: capi.call(f"lafem_hip_dot_{tag}", sp, x, y, n, res))
        keep.clear()
        torch.cuda.synchronize()
    # attainable HBM rates on this GPU: a pure read (the norm of a 2 GiB vector) and a device copy
    big = torch.empty(1 << 28, dtype=torch.float64, device="cuda").uniform_()
    big2 = torch.empty_like(big)
    res = torch.zeros(1, dtype=torch.float64, device="cuda")
    pb_, pb2, pr = ctypes.c_void_p(big.data_ptr()), ctypes.c_void_p(big2.data_ptr()), ctypes.c_void_p(res.data_ptr())
    nbig = big.numel()
    add("calibration: read (lafem_hip_norm2_f64, 2 GiB)", "calibration", "f64", 8 * nbig,
        lambda: capi.call("lafem_hip_norm2_f64", sp, pb_, nbig, pr))
    add("calibration: copy (feat3_hip_copy, 2 GiB)", "calibration", "f64", 16 * nbig,
        lambda: capi.call("feat3_hip_copy", sp, pb2, pb_, 8 * nbig))
    add("calibration: read (feat3_hip_read_probe, nontemporal, 2 GiB)", "calibration", "f64", 8 * nbig,
        lambda: capi.call("feat3_hip_read_probe", sp, pb_, 8 * nbig, pr))
    del big, big2
    spmv = [r for r in rows if r["format"] not in ("blas1", "calibration")]
    best = max(spmv, key=lambda r: r["GBs"])
    head = next(r for r in rows if r["kernel"] == "lafem_hip_apply_csr_f64_u32")
    roof = {"bound": "hbm", "achieved": head["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": head["frac_of_peak"],
            "traffic": None, "kernel": head["kernel"] + " (3D Q1 level 7, r = A x)", "bytes_per_launch": head["bytes"],
            "avg_launch_us": head["us"]}
    cfg = {"workload": f"drop-in Arch::Apply SpMV and BLAS-1 kernels, 3D Q1 level {lvl} Laplacian ({n} rows, {nnz} nnz) "
                       f"and the level-{max(lvl - 1, 1)} 3x3 vector Laplacian ({Av.rows} block rows)", "rows": n, "nnz": nnz}
    total_ms = sum(r["us"] for r in rows) * 1e-3
    return _common(args, head["GBs"], total_ms, "f64", "synthetic (assembled Laplacians, uniform[-1,1] vectors)", cfg, roof,
                   {"unit_note": "value = GB/s of the FP64 CSR apply; kernels[] lists every entry point",
                    "kernels": rows, "best_spmv": best["kernel"]}, None) | {"unit": "GB/s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=None,
                    help="finest level (default 7 for poisson3d, 9 for poisson2d, 6 for vector32, 10 for coeff16)")
    ap.add_argument("--inner", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--fmt", default="banded", choices=["csr", "sell", "banded"],
                    help="level-matrix format of the V-cycle (banded falls back to SELL-C-sigma where a matrix is not banded)")
    ap.add_argument("--band-dict", type=int, default=1, choices=[0, 1],
                    help="value-coded bands where every diagonal has at most 16 distinct values (1, default) "
                         "or the values stored (0); the same results bit for bit")
    ap.add_argument("--numbering", default="feat", choices=["feat", "lex"],
                    help="DOF numbering of the input hierarchy: feat = the reference's (StandardRefinery + "
                         "Lagrange DOF mapping) with the lattice order as the solver's internal level order; "
                         "lex = lexicographic input")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP64 / FP16 comparison solves")
    ap.add_argument("--workload", default="poisson3d", choices=["poisson3d", "poisson2d", "vector32", "coeff16", "spmv"],
                    help="poisson3d: the headline 3D Q1 L7 solve; poisson2d: 2D Q2 L9 (1,050,625 DOF); vector32: 32 "
                         "right-hand sides of the 3x3-block vector Laplacian (3D Q1 L6) in one batched solve; coeff16: "
                         "forward + adjoint solves and dL/dk of 16 coefficient fields (2D Q1 L10); spmv: every drop-in "
                         "Arch::Apply SpMV and BLAS-1 kernel at full size (GB/s against the HBM peak)")
    args = ap.parse_args()
    if args.level is None:
        args.level = {"poisson3d": 7, "poisson2d": 9, "vector32": 6, "coeff16": 10, "spmv": 7}[args.workload]

    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(no multi-GPU partition of the hierarchy is implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from feat3_amd import problems as pb
    from feat3_amd.solver import MGPCG

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.workload in ("vector32", "coeff16", "spmv"):
        out = {"vector32": bench_vector32, "coeff16": bench_coeff16, "spmv": bench_spmv}[args.workload](args)
        print(json.dumps(out), flush=True)
        return
    dim, deg = (3, 1) if args.workload == "poisson3d" else (2, 2)
    shape = f"{dim}D Poisson Q{deg} level {args.level}"

    t = time.time()
    H = pb.hierarchy(dim, deg, args.level, 1, numbering=args.numbering)
    host_setup_s = time.time() - t
    b = pb.filtered_rhs(H, 0)
    n, nnz = H.n, H.levels[0].A.nnz
    log(f"setup: {len(H.levels)} levels, n={n}, nnz={nnz}, {args.numbering} numbering ({host_setup_s:.1f} s)")
    # the same setup on the device: assembly (bit-identical to the host's, renumbered on the device
    # for the FEAT numbering; the mesh's numbering itself, feat_order, is computed beforehand) and
    # the solver's init from the device levels (renumbering and formats on the device); the solves
    # below use the host path
    orders = H.orders if args.numbering == "feat" else None
    MGPCG(pb.hierarchy_device(dim, deg, 3, 1, numbering=args.numbering), inner=args.inner, fmt=args.fmt).init().done()
    torch.cuda.synchronize()
    t = time.perf_counter()
    levels_dev = pb.hierarchy_device(dim, deg, args.level, 1, numbering=args.numbering, orders=orders)
    torch.cuda.synchronize()
    dev_setup_ms = 1e3 * (time.perf_counter() - t)
    t = time.perf_counter()
    sdev = MGPCG(levels_dev, inner=args.inner, fmt=args.fmt).init()
    torch.cuda.synchronize()
    dev_init_ms = 1e3 * (time.perf_counter() - t)
    sdev.done()
    del levels_dev, sdev
    log(f"setup on the device: assembly {dev_setup_ms:.1f} ms, solver init {dev_init_ms:.0f} ms "
        f"(host: assembly {host_setup_s:.2f} s, init below)")

    setup = {"host_hierarchy_s": host_setup_s, "device_hierarchy_ms": dev_setup_ms, "device_init_ms": dev_init_ms,
             "numbering": args.numbering}
    db = torch.from_numpy(b).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)

    def run(solver, steps, warmup):
        for _ in range(warmup):
            x.zero_()
            solver.correct(x, db)
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

    def variant(inner, steps, warmup, fmt=None, cycle="v", level_orders="auto"):
        fmt = fmt or args.fmt
        t0 = time.time()
        s = MGPCG(H, inner=inner, fmt=fmt, cycle=cycle, level_orders=level_orders,
                  options={"band_dict": args.band_dict}).init()
        init_s = time.time() - t0
        if inner == args.inner and fmt == args.fmt and cycle == "v":
            setup["host_init_s"] = init_s
        log(f"init {inner}/{fmt}/{cycle}: {init_s:.1f} s")
        times, iters = run(s, steps, warmup)
        hist = s.history()
        sol = x.cpu().numpy().copy()
        sweep_ms, sweep_stored_bytes = s.time_kernel(0, 50)
        outer_ms, outer_bytes = s.time_kernel(2, 50)
        vc_ms, _ = s.time_kernel(3, 10)
        info = s.info()
        s.done()
        ms = 1e3 * sum(times) / len(times)
        # the value-stored equivalent (the entries' values) beside the bytes the format reads
        sb_values = sweep_algorithmic_bytes(dict(info, banded=1) if info["banded"] == 3 else info, n,
                                            band_nnz(H.levels[0].A, inner, upper=info["banded"] == 2),
                                            info["sell_chunks"])
        sb = sweep_stored_bytes if info["banded"] == 3 else sb_values
        r = dict(ms_per_solve=ms, ms_min=1e3 * min(times), iterations=iters,
                 dof_iter_per_s=n * sum(iters) / sum(times), final_rel_residual=float(hist[-1] / hist[0]),
                 sweep_us=sweep_ms * 1e3, sweep_bytes=sb, sweep_GBs=sb / (sweep_ms * 1e-3) / 1e9,
                 level_format={0: "csr/sell", 1: "band", 2: "symmetric band", 3: "value-coded band"}[info["banded"]],
                 sweep_value_bytes=sb_values, sweep_value_GBs=sb_values / (sweep_ms * 1e-3) / 1e9,
                 vcycle_ms=vc_ms, kernels_per_vcycle=info["kernels_per_vcycle"],
                 outer_product_us=outer_ms * 1e3, outer_product_bytes=outer_bytes,
                 outer_product_GBs=outer_bytes / (outer_ms * 1e-3) / 1e9)
        log(f"{inner}/{fmt}/{cycle}: {ms:.3f} ms/solve, it {iters}, rel.res {r['final_rel_residual']:.2e}, "
            f"sweep {r['sweep_us']:.1f} us {r['sweep_GBs']:.0f} GB/s, {cycle.upper()}-cycle {vc_ms:.3f} ms")
        return r, sol, times, iters

    head, xm, times, iters = variant(args.inner, args.steps, args.warmup)
    ms_per_step = head["ms_per_solve"]
    value = n * sum(iters) / sum(times)

    variants = {}
    if not args.no_variants:
        for inner in ("f64", "f32", "f16"):
            if inner == args.inner:
                continue
            r, xv, _, _ = variant(inner, max(2, args.steps // 2), 1)
            r["speedup_of_headline"] = r["ms_per_solve"] / ms_per_step
            variants[inner] = (r, xv)
        # the level-matrix formats at the headline precision (CSR vs SELL-C-sigma vs banded)
        for fmt in ("csr", "sell", "banded"):
            if fmt == args.fmt:
                continue
            r, _, _, _ = variant(args.inner, max(2, args.steps // 2), 1, fmt=fmt)
            variants[f"{args.inner}/{fmt}"] = (r, None)
        # FEAT-numbered input without the lattice order: the level matrices keep FEAT's numbering
        # (no band structure: SELL-C-sigma), what the solver does when the caller gives no order
        if args.numbering == "feat":
            r, _, _, _ = variant(args.inner, max(2, args.steps // 2), 1, level_orders=None)
            variants[f"{args.inner}/{args.fmt}/feat-order-kept"] = (r, None)
        # the reference's F- and W-cycles as the preconditioner (multigrid.hpp:1366-1519)
        for cycle in ("f", "w"):
            r, _, _, _ = variant(args.inner, max(2, args.steps // 2), 1, cycle=cycle)
            variants[f"{args.inner}/{args.fmt}/{cycle}-cycle"] = (r, None)
    fp64 = variants.get("f64")
    if fp64:
        x64 = fp64[1]
        head["rel_diff_vs_fp64_solution"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))

    # CPU baseline: the oracle's FP64 PCG-MG (FEAT's generic loops restated) on one host core,
    # one full solve of the same system, solve phase only
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        S = po.System.from_hierarchy(H)
        xo = np.zeros(n)
        st, it, _ = S.solve(xo, b, po.params())
        dt = po.lib().orc_last_solve_seconds()
        cpu = {"value": n * it / dt, "unit": "DOF*iter/s", "cores": 1, "kind": "port",
               "sample": f"one full FP64 PCG-MG solve of the same {shape} system on one core "
                         f"({it} iterations, {dt:.2f} s solve phase), oracle restating FEAT's generic CPU loops",
               "time_to_tol_ms": dt * 1e3}
        if fp64:
            cpu["rel_diff_gpu_fp64_vs_cpu"] = float(np.linalg.norm(fp64[1] - xo) / np.linalg.norm(xo))
        cpu["rel_diff_headline_vs_cpu_fp64"] = float(np.linalg.norm(xm - xo) / np.linalg.norm(xo))
        cpu["iterations_cpu"] = int(it)
        # the same restatement with its row and BLAS-1 loops parallelised (OpenMP, all host threads
        # the job may use); reported beside the single-core reference analogue
        So = po.System.from_hierarchy(H, omp=True)
        xp = np.zeros(n)
        _, itp, _ = So.solve(xp, b, po.params())
        dtp = po.lib(True).orc_last_solve_seconds()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu["all_cores"] = {"value": n * itp / dtp, "unit": "DOF*iter/s", "cores": threads,
                            "time_to_tol_ms": dtp * 1e3, "iterations": int(itp)}
        cpu["host"] = _host_cpu()
        log(f"cpu baseline: {cpu}")

    achieved = head["sweep_GBs"]
    # measured with the value-coded band (the default)
    traffic = MEASURED_SWEEP_TRAFFIC.get((args.workload, args.level, args.inner, args.fmt)) if args.band_dict else None
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
        "dtype": {"f32": "f32", "f64": "f64", "f16": "f16-values/f32"}[args.inner],
        "data": f"synthetic (assembled {dim}D Q{deg} Poisson system, sine-bubble load, "
                f"{'FEAT' if args.numbering == 'feat' else 'lexicographic'} DOF numbering)",
        "config": {"workload": f"{shape} ({n} DOF, {nnz} nnz), PCG(FP64) + MG V(4,4) "
                               f"Richardson/Jacobi 0.5, exact coarse {'2x2x2' if dim == 3 else '2x2'} cells, {args.inner} V-cycle, "
                               f"{args.fmt} format, rel. tol 1e-8",
                   "dofs": n, "nnz": nnz, "levels": len(H.levels)},
        "time_to_tol_ms": ms_per_step,
        "iterations": float(np.mean(iters)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sweep, finest level (fused damped-Jacobi step)",
                     "bytes_per_launch": head["sweep_bytes"], "avg_launch_us": head["sweep_us"]},
        "headline": head,
        "variants": {k: v[0] for k, v in variants.items()},
        "speedup_vs_fp64": (fp64[0]["ms_per_solve"] / ms_per_step) if fp64 else None,
        "setup": setup,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
