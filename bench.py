"""Benchmark: batched weighted solve_batch on one MI355X (config 2 of the hot path).

Workload (default): n = 5,000,000 observations, factors A (200,000 levels, Zipf(1.1)
frequencies), B (10,000 uniform levels), C (50 uniform levels); weights ~ U(0.5, 2);
k = 16 float64 right-hand sides ~ N(0,1); tol 1e-8; default additive Schwarz.  A "step" is
one solve_batch of the 16 right-hand sides from inputs resident in HBM.  The headline value
is the mixed-precision form (FP32 preconditioner coefficients, FP64 Krylov); the plain
FP64 form is reported beside it.

Prints ONE JSON line.  The CPU baseline is the CPU oracle (a C restatement of the
reference's LSMR + additive Schwarz) solving a bounded sample of the same workload on
the host's cores with the same preconditioner data.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per k_gather_wave dispatch at the default workload (n = 5,000,000, 3 factors,
# k = 16), measured with rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a 1 GiB copy of the
# same run (profiles/k_gather_wave_n5M_k16_hbm_traffic.txt); None for other workloads.
GATHER_TRAFFIC_BYTES = {(5_000_000, 3, 16): 0.813631e9 + 0.674607e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cfg2_data(n=5_000_000, k=16, seed=2):
    rng = np.random.default_rng(seed)
    p = (np.arange(200_000) + 1.0) ** -1.1
    p /= p.sum()
    A = rng.choice(200_000, n, p=p)
    B = rng.integers(0, 10_000, n)
    Cc = rng.integers(0, 50, n)
    codes = np.asfortranarray(np.column_stack([A, B, Cc]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, k)))
    return codes, w, Y


def algorithmic_bytes(n, F, weighted, V, k, N):
    """Compulsory bytes of one D v gather (with the fused u update) and one D' reduction."""
    w = 1 if weighted else 0
    b_d = n * (4 * F + 8 * w + 8 * V + 16 * k) + 8 * k * N
    b_dt = n * (4 * F + 8 * w + 8 * V + 8 * k) + 8 * k * N
    return b_d, b_dt


def sorted_term_levels(codes):
    """Levels of the term the observations are ordered by: the largest, ties to the last."""
    levels = [int(codes[:, t].max()) + 1 for t in range(codes.shape[1])]
    best = max(levels)
    return levels[max(i for i, L in enumerate(levels) if L == best)]


def run_gpu(solver, Y_dev, k, n, N, precision, steps, warmup, tol):
    import torch
    from within_amd import _lib
    L = _lib.lib()
    opts = _lib.LsmrOpts()
    L.wx_lsmr_options_default(C.byref(opts))
    opts.tol = tol
    opts.precision = 1 if precision == "mixed" else 0
    X = torch.empty((k, N), dtype=torch.float64, device="cuda")   # column-major N x k
    Dm = torch.empty((k, n), dtype=torch.float64, device="cuda")  # column-major n x k
    stats = (_lib.RhsStats * k)()
    stream = torch.cuda.current_stream().cuda_stream

    def step(timing=False):
        # the timed solves do the reference's work (x, demeaned, the LSMR residual estimate); the
        # event-timed solve after them also forms the FP64 true residual that the line reports
        opts.timing = 1 if timing else 0
        opts.true_residual = 1 if timing else 0
        _lib.check(L.wx_solver_solve_batch_device(solver._h, Y_dev.data_ptr(), n, k, C.byref(opts), X.data_ptr(),
                                                  Dm.data_ptr(), stats, C.c_void_p(stream)))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # one more step with HIP-event timing of the phases (not part of the timed region)
    step(timing=True)
    tm = (C.c_double * 8)()
    L.wx_last_timing(tm, 8)
    phases = {name: (tm[2 * i], int(tm[2 * i + 1])) for i, name in
              enumerate(["gather", "adjoint", "precond", "iteration"])}
    st = [stats[j] for j in range(k)]
    return elapsed, st, phases, Dm


def run_plain_fp64(solver, Y_dev, n, N, tol, n_rhs):
    """Plain FP64 form: each right-hand side solved on its own (k = 1), FP64 coefficients."""
    import torch
    from within_amd import _lib
    L = _lib.lib()
    opts = _lib.LsmrOpts()
    L.wx_lsmr_options_default(C.byref(opts))
    opts.tol = tol
    opts.precision = 0
    opts.true_residual = 0
    X = torch.empty((1, N), dtype=torch.float64, device="cuda")
    Dm = torch.empty((1, n), dtype=torch.float64, device="cuda")
    stats = (_lib.RhsStats * 1)()
    stream = torch.cuda.current_stream().cuda_stream

    def one(j):
        ptr = Y_dev.data_ptr() + j * n * 8
        _lib.check(L.wx_solver_solve_batch_device(solver._h, ptr, n, 1, C.byref(opts), X.data_ptr(), Dm.data_ptr(),
                                                  stats, C.c_void_p(stream)))
        return stats[0].iterations

    one(0)  # warm-up (workspace of width 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = [one(j) for j in range(n_rhs)]
    torch.cuda.synchronize()
    return n_rhs / (time.perf_counter() - t0), its


def schwarz_bytes(pre_bytes: bytes, k: int, value_bytes: int, N: int) -> dict:
    """Algorithmic bytes of one additive Schwarz apply (the per-apply formula of the hot-path
    scope: factor and cross-tab entries at (index + value) bytes, each read twice -- once by each
    elimination SpMV, once by each triangular sweep --, dense kept blocks read once, counted as
    their nk (nk + 1) / 2 distinct entries since S^+ is symmetric, the slot-space vectors written
    and read four times, plus the FP64 DOF vectors: input and output, and the LSMR steps folded
    into the apply -- A'u read, p~ read and written by the restriction, p~ read again by the
    prolongation for the alpha dot products), from the preconditioner's exported arrays
    (little-endian WXPC layout)."""
    import struct
    assert pre_bytes[:4] == b"WXPC"
    _, kind, n = struct.unpack_from("<IIQ", pre_bytes, 4)
    if kind != 1:
        return {}
    sizes = {}
    pos = 20
    fields = [("slot_start", 8), ("ne", 4), ("nk", 4), ("fac_kind", 1), ("gauge", 1), ("glob", 4), ("pw", 8),
              ("fold", 8), ("inv_de", 8), ("c_ptr", 8), ("c_idx", 4), ("c_val", 8), ("dense_off", 8), ("dense", 8),
              ("l_row_off", 8), ("l_ptr", 8), ("l_idx", 4), ("l_val", 8), ("piv_off", 8), ("piv", 8),
              ("cover_off", 8), ("cover_perm", 4)]
    for name, sz in fields:
        m, = struct.unpack_from("<Q", pre_bytes, pos)
        sizes[name] = m
        pos += 8 + m * sz
    s = value_bytes
    slots = sizes["glob"]
    # the kept-block sizes of the dense components (fac_kind 0), from the arrays' positions
    nc = sizes["ne"]
    pos = 20 + 8 + sizes["slot_start"] * 8 + 8 + nc * 4
    nk = struct.unpack_from(f"<{nc}I", pre_bytes, pos + 8)
    kinds = pre_bytes[pos + 8 + nc * 4 + 8: pos + 8 + nc * 4 + 8 + nc]
    dense_sym = sum(n * (n + 1) // 2 for n, kd in zip(nk, kinds) if kd == 0)
    b_c = sizes["c_idx"] * (4 + s) * 2
    b_dense = dense_sym * s
    b_l = sizes["l_idx"] * (4 + s) * 2
    b_vec = s * k * slots * 4 + 8 * k * N * 2 + 8 * k * N * 4
    return {"bytes": b_c + b_dense + b_l + b_vec, "cross_tab_entries": sizes["c_idx"], "dense_entries": dense_sym,
            "factor_entries": sizes["l_idx"], "slots": slots}


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process may use, the thread
    share it is given (OMP_NUM_THREADS, 16 per GPU on the GPU box) and the CPU model (as lscpu
    prints it, from /proc/cpuinfo)."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", usable) or usable)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": nproc, "usable_cpus": usable, "thread_share": min(share, usable), "model": model}


def cpu_baseline(codes, w, Y, Dm_gpu, reps=5):
    """The CPU oracle (a C restatement of the reference's mLSMR + additive Schwarz) on the whole
    workload: all k right-hand sides at full n, solved right-hand-side parallel on the host's thread
    share (one column per thread, the reference's solve_batch decomposition, solver.rs:541-544),
    with the reference's default local solver (exact dense factor only for kept blocks of at most
    24 levels, sampled Schur + approximate Cholesky otherwise).  Best of `reps` solves
    (crates/within/benches/wallclock_1m3fe.rs:40-47); the preconditioner set-up is timed
    separately.  Also returns how far the GPU's demeaned output is from this CPU solve."""
    from oracle.oracle import OracleLib
    from tests._helpers import oracle_schwarz_from_bytes
    import within_amd as wa
    from within_amd._within import build_preconditioner
    lib = OracleLib()
    host = host_info()
    k = Y.shape[1]
    threads = max(1, min(host["thread_share"], k))
    t0 = time.perf_counter()
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=24, dense_policy="reference"))
    p = build_preconditioner(codes, w, cfg)
    sch = oracle_schwarz_from_bytes(lib, p.to_bytes())
    d = lib.design(codes)
    setup = time.perf_counter() - t0
    times = []
    Dm = res = None
    for rep in range(reps):
        t0 = time.perf_counter()
        _, Dm, res = lib.solve_batch(d, Y, w, prec="schwarz", schwarz=sch, tol=1e-8, threads=threads)
        times.append(time.perf_counter() - t0)
        log(f"[bench] cpu baseline solve {rep + 1}/{reps}: {times[-1]:.2f} s")
    best = min(times)
    rel = [float(np.linalg.norm(Dm_gpu[:, j] - Dm[:, j]) / np.linalg.norm(Dm[:, j])) for j in range(k)]
    out = {"value": k / best, "unit": "RHS/s", "cores": threads, "kind": "port",
           "sample": f"the whole workload: all {k} right-hand sides at full n, one column per thread on {threads} "
                     f"threads, additive Schwarz with the reference's default local-solver rule (sampled Schur + "
                     f"approximate Cholesky above 24 kept levels), FP64, best of {reps} solves "
                     f"({', '.join(f'{t:.2f}' for t in times)} s), {max(r.iterations for r in res)} max iterations",
           "setup_s": setup, "host": host, "all_converged": all(r.converged for r in res)}
    return out, max(rel)


def bench_many(args):
    """Uneven batch of independent problems: 64 problems with n log-uniform in [10^4, 10^6],
    2 or 3 factors (n/20, n/200, 30 uniform levels), 1-8 RHS each, tol 1e-8, default additive
    Schwarz.  A step is one solve_many call over all problems from inputs resident in HBM;
    beside it, the same problems solved one solve_batch call after another on the same GPU."""
    import torch
    import within_amd as wa
    from within_amd import _lib
    from tests._configs import cfg4_problems
    L = _lib.lib()
    probs = cfg4_problems(seed=4, n_problems=args.problems)
    t0 = time.perf_counter()
    solvers = [wa.Solver(codes) for codes, _ in probs]
    setup = time.perf_counter() - t0
    P = len(probs)
    ks = [Y.shape[1] for _, Y in probs]
    ns = [Y.shape[0] for _, Y in probs]
    log(f"[bench] {P} problems, n {min(ns)}..{max(ns)} (total {sum(ns)}), {sum(ks)} RHS, setup {setup:.1f} s")
    Yd = [torch.from_numpy(np.ascontiguousarray(Y.T)).cuda() for _, Y in probs]
    Xd = [torch.empty((k, s.n_dofs), dtype=torch.float64, device="cuda") for k, s in zip(ks, solvers)]
    Dd = [torch.empty((k, n), dtype=torch.float64, device="cuda") for k, n in zip(ks, ns)]
    stats = [(_lib.RhsStats * k)() for k in ks]
    opts = _lib.LsmrOpts()
    L.wx_lsmr_options_default(C.byref(opts))
    opts.tol = args.tol
    opts.precision = 1
    stream = torch.cuda.current_stream().cuda_stream
    hs = (C.c_void_p * P)(*[s._h for s in solvers])
    yp = (C.c_void_p * P)(*[t.data_ptr() for t in Yd])
    ld = (C.c_uint64 * P)(*ns)
    kk = (C.c_uint64 * P)(*ks)
    xp = (C.c_void_p * P)(*[t.data_ptr() for t in Xd])
    dp = (C.c_void_p * P)(*[t.data_ptr() for t in Dd])
    sp = (C.c_void_p * P)(*[C.cast(st, C.c_void_p) for st in stats])

    def many():
        _lib.check(L.wx_solve_many_device(hs, P, yp, ld, kk, C.byref(opts), xp, dp, sp, C.c_void_p(stream)))

    def sequential():
        for p in range(P):
            _lib.check(L.wx_solver_solve_batch_device(solvers[p]._h, yp[p], ns[p], ks[p], C.byref(opts), xp[p], dp[p],
                                                      stats[p], C.c_void_p(stream)))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    if args.shape:
        wa.set_many_shape(*[int(v) for v in args.shape.split(",")])
    t_seq = timed(sequential) if not args.no_seq else float("nan")
    its_seq = [[stats[p][j].iterations for j in range(ks[p])] for p in range(P)] if not args.no_seq else None
    sweep = {}
    if args.sweep:
        for shp in [(1, 1, False, False), (8, 2, True, False), (1, 1, False, True), (2, 1, True, True),
                    (4, 1, True, True), (4, 2, True, True), (8, 2, True, True), (4, 4, True, True),
                    (8, 2, True, True, 32), (8, 2, True, True, 128), (8, 2, True, True, 256)]:
            wa.set_many_shape(*shp)
            sweep[str(shp)] = timed(many) * 1e3
            log(f"[bench] shape {shp}: {sweep[str(shp)]:.1f} ms")
        wa.set_many_shape(0)
    t_many = timed(many)  # the first warm-up call runs the autotuner
    its = [[stats[p][j].iterations for j in range(ks[p])] for p in range(P)]
    shape = wa.last_many_shape()
    # per-problem latency of the last call: time from the call's start to the problem's outputs
    lat = np.array([sum(stats[p][j].time_solve for j in range(ks[p])) for p in range(P)]) * 1e3
    conv = all(stats[p][j].converged for p in range(P) for j in range(ks[p]))
    tr = max(stats[p][j].true_residual for p in range(P) for j in range(ks[p]))
    log(f"[bench] solve_many {t_many * 1e3:.1f} ms, sequential {t_seq * 1e3:.1f} ms, shape {shape}")
    out = {
        "metric": "RHS solved/s to tol 1e-8, many independent problems in one call",
        "value": sum(ks) / t_many,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_many * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 Krylov + f32 preconditioner coefficients",
        "data": "synthetic (seeded generator of the uneven many-problem workload)",
        "config": {"workload": f"{P} independent problems, n log-uniform in [1e4, 1e6], 2-3 factors "
                               "(n/20, n/200, 30 uniform levels), 1-8 RHS each, tol 1e-8",
                   "problems": P, "total_rhs": sum(ks), "total_obs": sum(ns), "precision": "mixed"},
        "sequential_solve_batch_value": None if args.no_seq else sum(ks) / t_seq,
        "many_over_sequential": None if args.no_seq else t_seq / t_many,
        "launch_shape": shape,
        "shape_sweep_ms": sweep or None,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p95": float(np.percentile(lat, 95)),
                       "max": float(lat.max()), "min": float(lat.min())},
        "converged": conv,
        "true_residual_max": tr,
        "iterations_equal_sequential": (its == its_seq) if its_seq is not None else None,
        "max_iterations": max(max(i) for i in its),
        "setup_s": setup,
    }
    print(json.dumps(out), flush=True)


def bench_panel(args):
    """Worker-firm panel: 400,000 workers x 5 periods (n = 2,000,000), firm drawn Zipf(1.2) over
    50,000 firms with a 5 % move probability per period, 5 years; effects worker, firm with an
    intercept and one N(0,1) slope, year; k = 4 RHS; tol 1e-10; default additive Schwarz (slope
    pairs through the SDDM conversion and Gremban covers).  A step is one solve_batch of the 4
    right-hand sides from inputs resident in HBM."""
    import torch
    import within_amd as wa
    from within_amd import _lib
    from tests._configs import cfg3_data
    L = _lib.lib()
    worker, firm, year, z, Y = cfg3_data()
    n, k = Y.shape
    t0 = time.perf_counter()
    solver = wa.Solver([wa.Effect(worker, True), wa.Effect(firm, True, [z]), wa.Effect(year, True)])
    setup = time.perf_counter() - t0
    Y_dev = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    Dm = torch.empty((k, n), dtype=torch.float64, device="cuda")
    stats = (_lib.RhsStats * k)()
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for precision in ("mixed", "fp64"):
        opts = _lib.LsmrOpts()
        L.wx_lsmr_options_default(C.byref(opts))
        opts.tol = 1e-10
        opts.precision = 1 if precision == "mixed" else 0

        def step():
            _lib.check(L.wx_solver_solve_batch_device(solver._h, Y_dev.data_ptr(), n, k, C.byref(opts), None,
                                                      Dm.data_ptr(), stats, C.c_void_p(stream)))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        out[precision] = ((time.perf_counter() - t) / args.steps, [stats[j] for j in range(k)])
    el, st = out["mixed"]
    el64, st64 = out["fp64"]
    log(f"[bench] panel: mixed {el * 1e3:.1f} ms, fp64 {el64 * 1e3:.1f} ms, iterations {[s.iterations for s in st]}")
    print(json.dumps({
        "metric": "RHS solved/s to tol 1e-10, worker-firm panel with a firm slope",
        "value": k / el, "unit": "RHS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 Krylov + f32 preconditioner coefficients",
        "data": "synthetic (seeded generator of the worker-firm panel workload)",
        "config": {"workload": "worker-firm panel: 400,000 workers x 5 periods, 50,000 firms Zipf(1.2), move 5 %/period, "
                               "firm intercept + slope, year; k=4; tol 1e-10", "n": n, "k": k,
                   "n_dofs": solver.n_dofs, "precision": "mixed"},
        "fp64_value": k / el64, "iterations": [s.iterations for s in st], "fp64_iterations": [s.iterations for s in st64],
        "converged": all(s.converged for s in st), "true_residual_max": max(s.true_residual for s in st),
        "setup_s": setup,
    }), flush=True)


def bench_fwl(args):
    """The Frisch-Waugh-Lovell step after demeaning: G = A'A of a demeaned block A = [y~, X~]
    of n = 5,000,000 rows and 16 columns (a response and 15 regressors), column-major in HBM.
    A step is one wx_gram_device call (one pass over A plus the fixed-order combine of the
    workgroups' partial triangles, synchronised)."""
    import torch
    from within_amd import _lib
    L = _lib.lib()
    n, m = 5_000_000, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    A = torch.randn((m, n), dtype=torch.float64, device="cuda", generator=g)  # column-major n x m
    G = torch.empty((m, m), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        _lib.check(L.wx_gram_device(A.data_ptr(), n, m, n, G.data_ptr(), C.c_void_p(stream)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ref = (A @ A.t()).cpu().numpy()
    err = float(np.abs(G.cpu().numpy() - ref).max() / np.abs(ref).max())
    bytes_per = 8.0 * n * m
    log(f"[bench] fwl gram: {dt * 1e6:.1f} us per call, {bytes_per / dt / 1e9:.0f} GB/s, max rel diff vs torch {err:.1e}")
    print(json.dumps({
        "metric": "Gram matrix A'A of the demeaned block (FWL step) per second", "value": 1.0 / dt, "unit": "calls/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded N(0,1) block)",
        "config": {"workload": "FWL Gram: n=5,000,000 rows x 16 columns, column-major in HBM", "n": n, "m": m},
        "roofline": {"bound": "hbm", "achieved": bytes_per / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_per / dt / 1e9 / HBM_PEAK_GBS, "traffic": None, "kernel": "k_gram + k_gram_reduce",
                     "algorithmic_bytes_per_launch": bytes_per, "avg_launch_ms": dt * 1e3,
                     "note": "per call including the synchronisation; A is read once"},
        "max_rel_diff_vs_torch": err,
    }), flush=True)


def bench_grad(args):
    """Differentiable weighted projection: n = 1,000,000, factors of 50,000 Zipf(1.1) and 2,000
    uniform levels, k = 8, weights exp(N(0, 0.5^2)), loss sum(mask * demeaned^2).  A step is one
    forward (Solver for the weights: host preconditioner build and upload, then the device
    solve) and one backward (adjoint solve with the same Solver, fused gradient kernel)."""
    import torch
    from within_amd.autograd import project
    import within_amd as wa
    from tests._configs import cfg5_data
    codes, w, Y, mask = cfg5_data()
    Z = torch.tensor(Y, device="cuda", requires_grad=True)
    wt = torch.tensor(w, device="cuda", requires_grad=True)
    M = torch.tensor(mask, device="cuda")
    opts = wa.LsmrOptions(tol=1e-8)
    design, reuse = codes, None

    def step():
        Z.grad = None
        wt.grad = None
        t0 = time.perf_counter()
        R = project(Z, wt, design, opts, reuse)
        loss = (M * R * R).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return t1 - t0, time.perf_counter() - t1

    def timed():
        for _ in range(args.warmup):
            step()
        fw = bw = 0.0
        for _ in range(args.steps):
            a, b = step()
            fw += a
            bw += b
        return fw / args.steps, bw / args.steps

    fw, bw = timed()
    # a training loop can keep the preconditioner of earlier weights: still an SPD
    # preconditioner of the new system (only the iteration count can change)
    base = wa.Solver(codes, weights=w)
    reuse = base.preconditioner
    fw_r, bw_r = timed()
    # ... and the Solver of earlier weights, reweighted (device design copied, not rebuilt)
    design = base
    fw_s, bw_s = timed()
    log(f"[bench] grad: forward {fw * 1e3:.1f} ms, backward {bw * 1e3:.1f} ms; with a kept preconditioner "
        f"forward {fw_r * 1e3:.1f} ms, backward {bw_r * 1e3:.1f} ms; reweighting a kept Solver "
        f"forward {fw_s * 1e3:.1f} ms, backward {bw_s * 1e3:.1f} ms")
    print(json.dumps({
        "metric": "weighted projection forward+backward per second (differentiable in Z and w)",
        "value": 1.0 / (fw + bw), "unit": "fwd+bwd/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (fw + bw) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded generator of the differentiable-projection workload)",
        "config": {"workload": "n=1,000,000; A 50,000 Zipf(1.1), B 2,000 uniform; k=8; w=exp(N(0,0.5^2)); "
                               "loss sum(mask * demeaned^2)", "n": len(w), "k": Y.shape[1]},
        "forward_ms": fw * 1e3, "backward_ms": bw * 1e3,
        "note": "the forward builds the preconditioner for the new weights on the host each step",
        "kept_preconditioner": {"value": 1.0 / (fw_r + bw_r), "forward_ms": fw_r * 1e3, "backward_ms": bw_r * 1e3},
        "reweighted_solver": {"value": 1.0 / (fw_s + bw_s), "forward_ms": fw_s * 1e3, "backward_ms": bw_s * 1e3},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=5_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--workload", choices=["solve_batch", "many", "panel", "grad", "fwl"], default="solve_batch",
                    help="solve_batch: the weighted 5M-row batch (default); many: 64 uneven independent problems; "
                         "panel: worker-firm panel with a firm slope; grad: differentiable projection; "
                         "fwl: the Gram matrix of a demeaned block")
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--sweep", action="store_true", help="many: also time every launch shape")
    ap.add_argument("--shape", default="", help="many: fixed launch shape 'streams,iterations_per_graph,graph,pack[,tile]'")
    ap.add_argument("--no-seq", action="store_true", help="many: skip the one-after-another comparison")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported by this build "
              "(the solve batches are independent; see the sharding note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "many":
        bench_many(args)
        return
    if args.workload == "panel":
        bench_panel(args)
        return
    if args.workload == "grad":
        bench_grad(args)
        return
    if args.workload == "fwl":
        bench_fwl(args)
        return

    import torch
    import within_amd as wa
    from within_amd import _lib
    n_dev, name = _lib.device_info()
    log(f"[bench] device {name}; generating data n={args.n} k={args.k}")
    codes, w, Y = cfg2_data(args.n, args.k)
    # the first Solver of a process also pays the runtime's and libraries' initialisation; the
    # second one's set-up is the steady-state figure
    t0 = time.perf_counter()
    solver = wa.Solver(codes, weights=w)
    setup_first = time.perf_counter() - t0
    del solver
    t0 = time.perf_counter()
    solver = wa.Solver(codes, weights=w)
    setup = time.perf_counter() - t0
    n, k, N = args.n, args.k, solver.n_dofs
    log(f"[bench] setup (design + preconditioner build + upload) {setup:.2f} s "
        f"(first in the process {setup_first:.2f} s), n_dofs={N}")
    Y_dev = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()  # (k, n) row-major == (n, k) column-major

    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy probe: 1 GiB read + 1 GiB written")
        del src, dst

    res = {}
    dm_mixed = None
    for precision in (["mixed"] if args.no_fp64 else ["mixed", "fp64"]):
        el, st, ph, Dm = run_gpu(solver, Y_dev, k, n, N, precision, args.steps, args.warmup, args.tol)
        res[precision] = (el, st, ph)
        if precision == "mixed":
            dm_mixed = Dm.cpu().numpy().T  # (n, k) column-major
        log(f"[bench] {precision}: {el / args.steps * 1e3:.1f} ms/step, iterations {[s.iterations for s in st]}, "
            f"max true residual {max(s.true_residual for s in st):.2e}, phases {ph}")
    el, st, ph = res["mixed"]
    value = k * args.steps / el
    F = codes.shape[1]
    b_d, b_dt = algorithmic_bytes(n, F, True, 0, k, N)
    # the timed solve launches one gather per iteration plus one after the last column has
    # converged, which returns at once; the per-launch duration is over the working launches
    iters = max(s.iterations for s in st)
    g_ms, _ = ph["gather"]
    a_ms, _ = ph["adjoint"]
    n_sorted = sorted_term_levels(codes)
    # k_gather_wave: the gather with the u update (B_D) plus the sorted term's D' output
    kern = "k_gather_wave (D v + u update + D' of the sorted term; a wavefront per item, two columns per lane)"
    bytes_per = b_d + 8 * k * n_sorted
    dur = g_ms / iters * 1e-3
    achieved = bytes_per / dur / 1e9
    # the remaining terms' D' (read u, w and codes once; write their levels)
    b_rest = n * (8 * k + 8 + 4 * (F - 1)) + 8 * k * (N - n_sorted)
    adj_dur = a_ms / iters * 1e-3
    out = {
        "metric": "RHS solved/s to tol 1e-8 (batched mixed-prec vs plain FP64 GPU); D/D' HBM GB/s",
        "value": value,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 Krylov + f32 preconditioner coefficients",
        "data": "synthetic (seeded generator of the weighted solve_batch workload)",
        "config": {"workload": "weighted solve_batch: n=5,000,000; A 200,000 Zipf(1.1), B 10,000 uniform, "
                               "C 50 uniform; k=16; weights U(0.5,2); tol 1e-8",
                   "n": n, "k": k, "n_dofs": N, "precision": "mixed"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": GATHER_TRAFFIC_BYTES.get((n, F, k)), "kernel": kern,
                     "algorithmic_bytes_per_launch": bytes_per, "avg_launch_ms": dur * 1e3},
        "roofline_adjoint_rest": {"achieved": b_rest / adj_dur / 1e9, "unit": "GB/s",
                                  "frac": b_rest / adj_dur / 1e9 / HBM_PEAK_GBS,
                                  "algorithmic_bytes_per_iteration": b_rest, "ms_per_iteration": adj_dur * 1e3,
                                  "what": "D' of the other terms (fused pair pass, piece and split-level sums)"},
        "iterations": [s.iterations for s in st],
        "true_residual_max": max(s.true_residual for s in st),
        "setup_s": setup,
        "setup_first_in_process_s": setup_first,
        "phases_ms": {p: {"total_ms": v[0], "launches": v[1]} for p, v in ph.items()},
    }
    pre = solver.preconditioner
    if pre is not None:
        sb = schwarz_bytes(pre.to_bytes(), k, 4, N)
        p_ms, p_launches = ph["precond"]
        if sb and p_launches:
            # as for the gather, the launch after the last column converged returns at once
            per = p_ms / iters * 1e-3
            out["roofline_precond"] = {"achieved": sb["bytes"] / per / 1e9, "unit": "GB/s",
                                       "frac": sb["bytes"] / per / 1e9 / HBM_PEAK_GBS,
                                       "algorithmic_bytes_per_apply": sb["bytes"], "ms_per_apply": per * 1e3,
                                       "what": "additive Schwarz apply (mixed: FP32 coefficients and slot vectors)",
                                       **{kk: vv for kk, vv in sb.items() if kk != "bytes"}}
    if "fp64" in res:
        el64, st64, ph64 = res["fp64"]
        out["fp64_value"] = k * args.steps / el64
        out["mixed_over_fp64"] = out["value"] / out["fp64_value"]
        out["fp64_iterations"] = [s.iterations for s in st64]
    if not args.no_plain:
        pv, pits = run_plain_fp64(solver, Y_dev, n, N, args.tol, k)
        out["plain_fp64_value"] = pv
        out["plain_fp64_iterations"] = pits
        out["mixed_over_plain_fp64"] = out["value"] / pv
        log(f"[bench] plain fp64 (one RHS per solve): {pv:.1f} RHS/s")
    if not args.no_cpu:
        try:
            out["cpu_baseline"], out["max_rel_demeaned_vs_cpu"] = cpu_baseline(codes, w, Y, dm_mixed)
            log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.2f} RHS/s; GPU (mixed) demeaned vs the CPU "
                f"solve: max relative difference {out['max_rel_demeaned_vs_cpu']:.2e}")
        except Exception as e:  # the baseline is reported, never the measured value
            log(f"[bench] cpu baseline failed: {e!r}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
