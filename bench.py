"""Benchmark of the numeric phase on one MI355X.

Workload (one "step" = refactorization + solve with 8 right-hand sides):
3-D 7-point Laplacian on a 64^3 grid (n = 262,144), SPD, Float64 values,
nested dissection of the matrix graph (reordering_alg "algo4"), amalgamation (512, 0.25, 8), B ~ N(0, 1) n x 8.
Two forms are timed: the native FP64 factor + solve, and the FP32 factor with
FP64 iterative refinement (the residual and the update of x in FP64) run for
the number of steps that reaches a normwise backward error <= 1e-10.

`value` is the FP64 form's refactorize+solve milliseconds per step (lower is
better); the FP32+IR form, factor TFLOP/s, backward errors, the dominant
kernel's roofline (HIP events inside the library, on the solver's stream) and
a CPU baseline (the reference's CPU path with LAPACK/BLAS per front on the whole
workload, see `cpu_baseline.sample`) are on the same JSON line, with the GPU
factors' largest element-wise errors against that CPU path's FP64 factor.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sparsedirectsolver.jl_amd"))

import numpy as np  # noqa: E402

PEAK = {"fp64_tflops": 78.6, "fp32_tflops": 157.3, "hbm_gbs": 8000.0}
MFMA_KINDS = ("update_panel", "update_cb", "front_small", "potrf", "trsm", "pivot_panel", "inverse")


# launch class of the library's profile -> the kernel it times
KIND_KERNEL = {"trsm": "k_big_step", "update_cb": "k_big_update_mfma", "extend_add": "k_big_extend_add"}


def measured_traffic(kind):
    """HBM bytes per launch of the class's kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries (copy-calibrated, profiles/*_hbm_traffic.txt), or None when none holds a
    figure for that kernel."""
    import glob
    name = KIND_KERNEL.get(kind)
    if name is None:
        return None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.txt"))):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == name and f[1] == "traffic_bytes_per_launch":
                return float(f[2])
    return None


def bench_copy(args):
    """Calibration of the HBM counters: a plain device copy of 512 MiB, args.steps times (the
    FETCH_SIZE / WRITE_SIZE reading of a copy of known size scales the other kernels' readings)."""
    import torch
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_per_copy": a.numel() * 8, "copies": args.steps}), flush=True)


def backward_error(A, X, B):
    """max over right-hand sides of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    R = B - A @ X
    an = abs(A).sum(axis=1).max()
    return float(np.max(np.abs(R).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))))


def make_workload(k, nrhs, seed=666):
    from sparsedirectsolver import matrices as M
    A = M.laplacian3d(k, k, k)
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    B = np.random.default_rng(seed).standard_normal((A.shape[0], nrhs))
    return A, (rp, cv, nz), None, B


TILE_KEYS = ("extend_add_rows", "cb_k_chunk", "trailing_split", "step_two_per_cu", "small_front_kernel",
             "subtree_kib", "panel_block", "pivot_panels")


def make_solver(csr, n, perm, amalgamation, factor_precision=None, tiles=None):
    """Analysis, the tile-shape autotuner on these values (or the given tile shapes), then the
    first factorization."""
    import sparsedirectsolver as sds
    rp, cv, nz = csr
    s = sds.DirectSolver((rp, cv, nz, n), "SPD", "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    else:
        sds.setparam(s, "reordering_alg", "algo4")  # nested dissection of the matrix graph
    sds.setparam(s, "amalgamation", amalgamation)
    if factor_precision:
        sds.setparam(s, "factor_precision", factor_precision)
    sds.execute("analysis", s)
    if tiles:
        s.tiles = dict(zip(TILE_KEYS, tiles))
        sds.set_tiles(s, s.tiles)
    else:
        s.tiles = sds.autotune(s, reps=2)
    sds.execute("factorization", s)
    return s


def solve_inverse_entries(S):
    """Entries of the diagonal-block inverses the Cholesky sweeps read: per front, per solve step of
    at most 512 columns, nw^2 (the inverse of the step's diagonal block: 64 x 64 blocks for the
    narrow steps, the whole nw x nw for the wide ones)."""
    tot = 0
    for w in S.front_ncols:
        w = int(w)
        for c0 in range(0, w, 512):
            nw = min(w, c0 + 512) - c0
            tot += nw * nw
    return tot


def time_steps(s, X, B, steps, warmup, ir_steps):
    """Median-free total timing of `steps` refactorization+solve steps (after `warmup`)."""
    import torch
    import sparsedirectsolver as sds
    sds.setparam(s, "ir_n_steps", ir_steps)
    for _ in range(warmup):
        sds.execute("refactorization", s)
        sds.execute("solve", s, X, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sds.execute("refactorization", s)
        sds.execute("solve", s, X, B)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def time_graph(s, X, B, steps, warmup):
    """Refactorization + solve (with the solver's refinement steps) captured once into a HIP
    graph and replayed: ms per replay, inputs resident (B is copied into the graph's buffer
    before the timed replays, X out after them)."""
    import torch
    from sparsedirectsolver.solver import graph_capture, graph_replay
    graph_capture(s, B.shape[1], "solve")
    graph_replay(s, None, B)
    for _ in range(warmup):
        graph_replay(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph_replay(s)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    graph_replay(s, X)
    torch.cuda.synchronize()
    return ms


def time_phase(s, steps, fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(S, nz, A, B):
    """The reference's CPU path (oracle/reference_path.py: ref_factorize! and ref_solve! of
    src/reference/cholesky.jl with LAPACK potrf and BLAS trsm / syrk / gemm per front) on the whole
    workload, timed on the host with one BLAS thread and with all the threads this process may
    use. `value` is the faster of the two (factorization + solve). Returns the result and the
    CPU factor (FP64), against which the GPU factors' errors are stated."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import reference_path
    from threadpoolctl import threadpool_limits
    cores = len(os.sched_getaffinity(0))
    runs = {}
    fac = None
    for th in (1, cores):
        with threadpool_limits(limits=th):
            t0 = time.perf_counter()
            info, f = reference_path.cholesky(S, nz)
            tf = time.perf_counter() - t0
            t0 = time.perf_counter()
            X = reference_path.solve(S, f, B)
            ts = time.perf_counter() - t0
        assert info == 0
        runs[th] = (tf, ts, backward_error(A, X, B))
        fac = f
        print(f"cpu baseline, {th} BLAS thread(s): factorization {tf:.2f} s, solve {ts:.2f} s", file=sys.stderr,
              flush=True)
    best = min(runs, key=lambda t: sum(runs[t][:2]))
    tf, ts, berr = runs[best]
    return {
        "value": round((tf + ts) * 1e3, 1), "unit": "ms", "cores": best, "kind": "port",
        "sample": (f"the whole workload: ref_factorize! + ref_solve! restated with LAPACK potrf / BLAS trsm, syrk "
                   f"per front (scipy's OpenBLAS) over the same analysis, {B.shape[1]} right-hand sides; the faster "
                   f"of 1 and {cores} BLAS threads"),
        "factor_ms": round(tf * 1e3, 1), "solve_ms": round(ts * 1e3, 1), "backward_error": berr,
        "factor_gflops": round(S.flops_true / tf / 1e9, 2),
        "by_threads": {str(t): {"factor_ms": round(v[0] * 1e3, 1), "solve_ms": round(v[1] * 1e3, 1)}
                       for t, v in runs.items()},
        "host": {"cpu": cpu_model(), "threads_available": cores},
    }, fac


def bench_batch(args):
    """The uneven batch: 1,024 2-D 5-point variable-coefficient Laplacians, sides log-uniform in
    [24, 140], conductivities exp(N(0, 1)), each with its own nested dissection; FP32 factor +
    FP64 refinement; refactorization + solve of the whole batch replayed from one HIP graph,
    against a loop over the members one at a time (each member its own solver, eager)."""
    import torch
    import sparsedirectsolver as sds
    from sparsedirectsolver import matrices as M
    from sparsedirectsolver.batch import UnevenBatch
    from sparsedirectsolver.solver import graph_capture, graph_replay
    rng = np.random.default_rng(666)
    count = args.members
    sides = np.round(np.exp(rng.uniform(np.log(24), np.log(140), count))).astype(int)
    mats = [M.lower_csr(M.variable_laplacian((int(q), int(q)), seed=1000 + k)) for k, q in enumerate(sides)]
    perms = [sds.grid_nd_perm(int(q), int(q), 1) for q in sides]
    amal = (128, 0.25, 8)
    t0 = time.perf_counter()
    batch = UnevenBatch(mats, "SPD", "L", perms=perms)
    batch.setparam("amalgamation", amal)
    batch.setparam("factor_precision", "Float32")
    batch.setparam("ir_n_steps", args.ir_steps)
    batch.execute("analysis")
    t_analysis = time.perf_counter() - t0
    s = batch.solver
    n = batch.n
    B = torch.tensor(rng.standard_normal((n, 1)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    print("batch: autotuning the tile shapes ...", file=sys.stderr, flush=True)
    tiles = sds.autotune(s, reps=1)
    print(f"batch: tiles {tiles}", file=sys.stderr, flush=True)
    sds.execute("factorization", s)
    graph_capture(s, nrhs=1)
    for _ in range(args.warmup):
        graph_replay(s, X, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph_replay(s, X, B)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) * 1e3 / args.steps
    A = batch.A + batch.A.T - __import__("scipy").sparse.diags(batch.A.diagonal())
    berr = backward_error(A, X.cpu().numpy(), B.cpu().numpy())
    # one system at a time
    solvers = []
    for m, p in zip(mats, perms):
        rp, cv, nz = M.csr_arrays(m)
        q = sds.DirectSolver((rp, cv, nz, m.shape[0]), "SPD", "L")
        sds.setparam(q, "user_perm", p)
        sds.setparam(q, "amalgamation", amal)
        sds.setparam(q, "factor_precision", "Float32")
        sds.setparam(q, "ir_n_steps", args.ir_steps)
        sds.execute("analysis", q)
        sds.execute("factorization", q)
        solvers.append(q)
    parts = batch.split(B)
    outs = [torch.zeros_like(b) for b in parts]
    def loop():
        for q, b, x in zip(solvers, parts, outs):
            sds.execute("refactorization", q)
            sds.execute("solve", q, x, b)
    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 4)):
        loop()
    torch.cuda.synchronize()
    t_loop = (time.perf_counter() - t0) * 1e3 / max(1, args.steps // 4)
    S = s.symbolic
    res = {
        "metric": "uneven batch refactor+solve ms (FP32 factor + FP64 IR, one graph) vs one-at-a-time loop",
        "value": round(t_batch, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batch, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64ir", "data": "synthetic",
        "config": {"workload": f"{count} 2-D 5-point variable-coefficient Laplacians, sides log-uniform in "
                               f"[24, 140] (sum n = {n}), 1 right-hand side, {args.ir_steps} refinement steps",
                   "ordering": "geometric nested dissection per member", "amalgamation": list(amal),
                   "nnz_stored": S.nnz_stored, "supernodes": S.ns, "factor_flops": S.flops_true,
                   "parallelism": "single GPU", "autotuned_tiles": tiles},
        "backward_error": berr, "loop_ms": round(t_loop, 3), "speedup_vs_loop": round(t_loop / t_batch, 2),
        "analysis_s": round(t_analysis, 2),
    }
    print(json.dumps(res), flush=True)


def bench_adjoint(args):
    """Differentiation (C5): the loss 0.5 ||x - x*||^2 with x = A^-1 b, b and x* ~ N(0, 1), on a
    48^3 variable-coefficient 7-point operator (log-normal conductivities, SPD Cholesky) and on
    the nonsymmetric upwind convection-diffusion variant (Peclet 10, LU with threshold partial
    pivoting), FP64 and FP32 factor + FP64 refinement. Gradients with respect to b and to the
    stored values of A come from the transposed solve on the same factor and the HIP scatter of
    -lambda x^T; they are checked against central finite differences on random stored entries
    (each difference is a refactorization + solve) and entries of b. Timed: refactorization +
    solve + transposed solve + scatter."""
    import torch
    import sparsedirectsolver as sds
    from sparsedirectsolver import matrices as M
    from sparsedirectsolver.adjoint import rrule_solve

    k0 = args.grid if args.grid != 64 else 48
    out = {"metric": "refactor+solve+adjoint ms, gradient vs central finite differences",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "data": "synthetic", "cases": []}
    rng = np.random.default_rng(11)
    for kind in ("spd", "lu"):
        k = k0 if kind == "spd" else args.lu_grid
        nfd = args.fd_entries if kind == "spd" else max(8, args.fd_entries // 4)
        if kind == "spd":
            A = M.variable_laplacian((k, k, k), seed=5)
            Ls = M.lower_csr(A)
            structure, view = "SPD", "L"
        else:
            A = M.convection_diffusion3d(k, peclet=10.0).tocsr()
            A.sort_indices()
            Ls = A
            structure, view = "G", "F"
        rp, cv, nz = M.csr_arrays(Ls)
        n = A.shape[0]
        bh, xstar = rng.standard_normal(n), rng.standard_normal(n)
        b = torch.tensor(bh, dtype=torch.float64, device="cuda")
        xs_t = torch.tensor(xstar, dtype=torch.float64, device="cuda")
        for fp in ("Float64", "Float32"):
            s = sds.DirectSolver((rp, cv, nz, n), structure, view)
            sds.setparam(s, "user_perm", sds.grid_nd_perm(k, k, k))
            sds.setparam(s, "amalgamation", (512, 0.25, 8))
            if fp == "Float32":
                sds.setparam(s, "factor_precision", "Float32")
                sds.setparam(s, "ir_n_steps", 2)
            sds.execute("analysis", s)
            t0 = time.perf_counter()
            sds.execute("factorization", s)
            torch.cuda.synchronize()
            print(f"{kind} {k}^3 {fp}: analysis + first factorization done ({time.perf_counter() - t0:.2f} s)",
                  file=sys.stderr, flush=True)

            def step():
                sds.execute("refactorization", s)
                x, pullback = rrule_solve(s, b)
                return x, pullback(x - xs_t)

            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                x, (bbar, abar) = step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            xh, bbar, abar = x.cpu().numpy(), bbar.cpu().numpy(), abar.cpu().numpy()
            Ah = A.tocsr()
            berr = float(np.abs(Ah @ xh - bh).max() / (np.abs(Ah).max() * np.abs(xh).max() + np.abs(bh).max()))
            print(f"  timed {ms:.2f} ms per step, backward error {berr:.2e}", file=sys.stderr, flush=True)

            def loss(zv, bv):
                sds.update(s, torch.tensor(zv, dtype=torch.float64, device="cuda"))
                sds.execute("refactorization", s)
                xx = torch.zeros_like(b)
                sds.execute("solve", s, xx, torch.tensor(bv, dtype=torch.float64, device="cuda"))
                return 0.5 * float(np.sum((xx.cpu().numpy() - xstar) ** 2))

            # central differences: truncation O(h^2), cancellation O(eps |loss| / h); the LU case's loss
            # is large, so it takes the larger step
            hrel = 1e-6 if fp == "Float64" and kind == "spd" else 1e-4
            picks = rng.choice(nz.shape[0], nfd, replace=False)
            gmax = max(np.abs(abar).max(), 1e-300)
            errs = []
            for p in picks:
                h = hrel * max(1.0, abs(nz[p]))
                zp, zm = nz.copy(), nz.copy()
                zp[p] += h
                zm[p] -= h
                errs.append(abs((loss(zp, bh) - loss(zm, bh)) / (2 * h) - abar[p]) / gmax)
                if len(errs) % 8 == 0:
                    print(f"  {len(errs)} finite differences, max rel err {max(errs):.2e}", file=sys.stderr, flush=True)
            berrs = []
            bmax = max(np.abs(bbar).max(), 1e-300)
            for i in rng.choice(n, 8, replace=False):
                h = hrel
                bp, bm = bh.copy(), bh.copy()
                bp[i] += h
                bm[i] -= h
                berrs.append(abs((loss(nz, bp) - loss(nz, bm)) / (2 * h) - bbar[i]) / bmax)
            loss(nz, bh)
            out["cases"].append({
                "matrix": (f"3-D variable-coefficient 7-point {k}^3, SPD Cholesky" if kind == "spd" else
                           f"3-D upwind convection-diffusion {k}^3 (Peclet 10), LU with threshold pivoting"),
                "n": n, "factor": fp, "ms_per_step": round(ms, 3), "backward_error": berr,
                "fd_entries": int(nfd), "fd_max_rel_err_A": float(max(errs)),
                "fd_max_rel_err_b": float(max(berrs)), "fd_step_rel": hrel})
            print(json.dumps(out["cases"][-1]), file=sys.stderr, flush=True)
            del s
    out["value"] = out["cases"][0]["ms_per_step"]
    out["unit"] = "ms"
    out["ms_per_step"] = out["value"]
    out["dtype"] = "f64"
    out["config"] = {"workload": f"differentiation of 0.5||x - x*||^2 through x = A^-1 b, {k0}^3 SPD and "
                                 f"{args.lu_grid}^3 LU",
                     "ordering": "geometric nested dissection", "amalgamation": [512, 0.25, 8]}
    print(json.dumps(out), flush=True)


def bench_kkt(args):
    """The KKT system [H J^T; J -delta I] (delta = 1e-8; H random sparse symmetric, about 10 entries
    per row, diagonal = row-abs-sum + U[0,1]; J with 4 entries per row at uniform random columns),
    LDL^T "S" with Bunch-Kaufman pivoting, threshold 0.01, AMD ordering, one right-hand side
    b ~ N(0, 1). Timed: refactorization + solve, FP64, and FP32 factor + FP64 refinement (plain and FGMRES). Checked:
    inertia (nh, nj, 0), and the FP64 pivot order, pivot kinds and D against the CPU oracle's
    LDL^T on the same matrix and analysis (also the CPU baseline). The random structure fills in
    like a dense matrix (nnz(L) grows about as n^2), so the default size is nh = 8000, nj = 4000."""
    import torch
    import sparsedirectsolver as sds
    from sparsedirectsolver import matrices as M
    from sparsedirectsolver.solver import numeric_pivots, numeric_diag

    nh = args.kkt_nh
    nj = nh // 2
    K = M.kkt_matrix(nh, nj)
    rp, cv, nz = M.csr_arrays(M.lower_csr(K))
    n = K.shape[0]
    bh = np.random.default_rng(666).standard_normal(n)
    b = torch.tensor(bh, dtype=torch.float64, device="cuda")
    amal = (512, 0.25, 8)
    cases = []
    s64 = None
    # the FP32 factor perturbs many tiny pivots: plain refinement stalls there, FGMRES (preconditioned
    # by the same factor, stopping at a relative residual of 1e-14) is the form meant for it
    for fp, mode in (("Float64", "ir"), ("Float32", "ir"), ("Float32", "fgmres")):
        s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
        sds.setparam(s, "amalgamation", amal)
        # AMD named explicitly: the automatic choice (fewest structurally zero pivots, then the cost
        # with its per-level term) takes nested dissection here, 1.6x the flops of AMD
        sds.setparam(s, "reordering_alg", "algo3")
        sds.setparam(s, "pivot_type", "A")
        sds.setparam(s, "pivot_threshold", 0.01)
        if fp == "Float32":
            sds.setparam(s, "factor_precision", "Float32")
        # delta = 1e-8 makes the system ill-conditioned: every form takes refinement steps
        sds.setparam(s, "ir_mode", mode)
        sds.setparam(s, "ir_n_steps", args.ir_steps if mode == "ir" else 20)
        if mode == "fgmres":
            sds.setparam(s, "ir_tol", 1e-14)
        sds.execute("analysis", s)
        t0 = time.perf_counter()
        sds.execute("factorization", s)
        torch.cuda.synchronize()
        print(f"kkt n={n} {fp} {mode}: analysis + first factorization done ({time.perf_counter() - t0:.2f} s)",
              file=sys.stderr, flush=True)
        x = torch.zeros_like(b)
        for _ in range(args.warmup):
            sds.execute("refactorization", s)
            sds.execute("solve", s, x, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sds.execute("refactorization", s)
            sds.execute("solve", s, x, b)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        t_ref = time_phase(s, args.steps, lambda: sds.execute("refactorization", s))
        xh = x.cpu().numpy()
        berr = backward_error(K, xh[:, None], bh[:, None])
        npos, nneg = sds.getparam(s, "inertia")
        cases.append({"factor": fp, "ir_mode": mode, "ir_steps": int(sds.getparam(s, "ir_n_steps")),
                      "ms_per_step": round(ms, 3), "refactor_ms": round(t_ref, 3),
                      "backward_error": berr, "inertia": [int(npos), int(nneg), int(n - npos - nneg)],
                      "inertia_expected": [nh, nj, 0], "npivots": int(sds.getparam(s, "npivots")),
                      "pivots_2x2": sds.getparam(s, "pivot_stats")["n2x2"]})
        if fp == "Float64":
            # the same refactorization + solve replayed from one HIP graph (launches pre-built)
            xg = torch.zeros(n, 1, dtype=torch.float64, device="cuda")
            cases[-1]["graph_ms_per_step"] = round(time_graph(s, xg, b.view(n, 1), args.steps, args.warmup), 3)
            cases[-1]["graph_backward_error"] = backward_error(K, xg.cpu().numpy(), bh[:, None])
        print(json.dumps(cases[-1]), file=sys.stderr, flush=True)
        if fp == "Float64":
            s64 = s
        else:
            del s
    S = s64.symbolic
    res = {"metric": "KKT LDL^T (Bunch-Kaufman) refactor+solve ms, inertia, FP64 pivots vs the CPU oracle",
           "value": cases[0]["ms_per_step"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": cases[0]["ms_per_step"], "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"KKT [H J^T; J -1e-8 I], H {nh} x {nh} (~10 nnz/row), J {nj} x {nh} (4 nnz/row), "
                                  f"n = {n}, LDL^T Bunch-Kaufman, threshold 0.01, one right-hand side",
                      "ordering": "AMD", "amalgamation": list(amal), "nnz_L": S.nnz_L, "nnz_stored": S.nnz_stored,
                      "supernodes": S.ns, "factor_flops": S.flops_true, "parallelism": "single GPU"},
           "cases": cases}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        print("cpu baseline: oracle LDL^T on the same analysis ...", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        F = oracle.ldlt(S, nz, pivot_type="A", pivot_threshold=0.01, eps=1e-13)
        t_cpu = (time.perf_counter() - t0) * 1e3
        piv, kind = numeric_pivots(s64)
        res["pivots_bitwise_vs_oracle"] = bool(np.array_equal(piv, F["piv"]) and np.array_equal(kind, F["kind"]))
        res["d_bitwise_vs_oracle"] = bool(np.array_equal(numeric_diag(s64), F["d"]))
        res["cpu_baseline"] = {"value": round(t_cpu, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the whole oracle LDL^T factorization of the same matrix and analysis "
                                         "(single thread)"}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-default-amalgamation", action="store_true",
                    help="laplacian3d: skip the step at the reference's default amalgamation (32, 0.25, 8)")
    ap.add_argument("--workload", choices=("laplacian3d", "batch", "adjoint", "kkt", "copy"), default="laplacian3d")
    ap.add_argument("--tiles", type=lambda t: [int(x) for x in t.split(",")], default=None,
                    help="laplacian3d: tile shapes instead of the autotuner (" + ",".join(TILE_KEYS) + ")")
    ap.add_argument("--kkt-nh", type=int, default=8000, help="kkt workload: order of H (J has half as many rows)")
    ap.add_argument("--fd-entries", type=int, default=64, help="adjoint workload: finite-difference checks")
    ap.add_argument("--lu-grid", type=int, default=48, help="adjoint workload: grid side of the LU case")
    ap.add_argument("--members", type=int, default=1024)
    ap.add_argument("--ir-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the numeric phase of one "
              "system runs on one MI355X, and multi-GPU sharding of the tree is not built).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "copy":
        bench_copy(args)
        return
    if args.workload == "kkt":
        bench_kkt(args)
        return
    import torch
    import sparsedirectsolver as sds
    from sparsedirectsolver.solver import numeric_factor, profile, profile_read

    amal = (512, 0.25, 8)
    A, csr, perm, Bh = make_workload(args.grid, args.nrhs)
    n = A.shape[0]
    B = torch.tensor(Bh, dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)

    # -- FP64 form
    s64 = make_solver(csr, n, perm, amal, tiles=args.tiles)
    S = s64.symbolic
    t_step64 = time_steps(s64, X, B, args.steps, args.warmup, 0)
    berr64 = backward_error(A, X.cpu().numpy(), Bh)
    t_graph64 = time_graph(s64, X, B, args.steps, args.warmup)
    berr64g = backward_error(A, X.cpu().numpy(), Bh)
    t_refact64 = time_phase(s64, args.steps, lambda: sds.execute("refactorization", s64))
    t_solve64 = time_phase(s64, args.steps, lambda: sds.execute("solve", s64, X, B))
    # profiled pass over the same steps: per-kernel-class HIP events on the solver's stream
    # (the events serialize the launches, so these phase times run longer than the ones above)
    profile(s64, True)
    time_phase(s64, args.steps, lambda: sds.execute("refactorization", s64))
    time_phase(s64, args.steps, lambda: sds.execute("solve", s64, X, B))
    prof = profile_read(s64)
    profile(s64, False)
    flops_true = S.flops_true
    res = {
        "metric": "refactor+solve ms & factor TFLOP/s, FP32+IR vs FP64 at ≤1e-10 backward err",
        "value": round(t_step64, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_step64, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3-D 7-point Laplacian {args.grid}^3 (n = {n}), SPD Cholesky, "
                               f"{args.nrhs} right-hand sides, refactorization + solve",
                   "ordering": "nested dissection of the matrix graph (reordering_alg algo4)",
                   "amalgamation": list(amal), "nrhs": args.nrhs,
                   "nnz_L": S.nnz_L, "nnz_stored": S.nnz_stored, "supernodes": S.ns,
                   "factor_flops": flops_true, "parallelism": "single GPU", "autotuned_tiles": s64.tiles},
        "fp64": {"refactor_ms": round(t_refact64, 3), "solve_ms": round(t_solve64, 3),
                 "factor_tflops": round(flops_true / (t_refact64 * 1e-3) / 1e12, 3),
                 "backward_error": berr64},
        "graph": {"fp64_ms_per_step": round(t_graph64, 3), "fp64_backward_error": berr64g},
    }
    # dominant kernel class of the profiled region and its roofline
    dom = max(prof, key=lambda k: prof[k][0])
    ms, nl, fl, by = prof[dom]
    if dom in MFMA_KINDS and fl > 0:
        ach = fl / (ms * 1e-3) / 1e12
        res["roofline"] = {"kernel": dom, "bound": "mfma", "achieved": round(ach, 3), "peak": PEAK["fp64_tflops"],
                           "unit": "TFLOP/s", "frac": round(ach / PEAK["fp64_tflops"], 4),
                           "traffic": measured_traffic(dom), "algorithmic_bytes_per_launch": round(by / nl),
                           "launches": nl, "avg_launch_ms": round(ms / nl, 4)}
    else:
        ach = by / (ms * 1e-3) / 1e9
        res["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": PEAK["hbm_gbs"],
                           "unit": "GB/s", "frac": round(ach / PEAK["hbm_gbs"], 4),
                           "traffic": measured_traffic(dom), "algorithmic_bytes_per_launch": round(by / nl),
                           "launches": nl, "avg_launch_ms": round(ms / nl, 4)}
    res["profile_ms"] = {k: round(v[0] / args.steps, 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    # the solve against the HBM roofline: its algorithmic bytes are the factor panels and the
    # diagonal-block inverses the sweeps multiply by (read once forward, once backward) and the
    # right-hand sides (B read, Y written and read per sweep, X written)
    ninv = solve_inverse_entries(S)
    sbytes = 8.0 * (2.0 * (S.nnz_stored + ninv) + 6.0 * n * args.nrhs)
    sach = sbytes / (t_solve64 * 1e-3) / 1e9
    res["solve_roofline"] = {"bound": "hbm", "achieved": round(sach, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                             "frac": round(sach / PEAK["hbm_gbs"], 4), "algorithmic_bytes": round(sbytes),
                             "factor_entries": S.nnz_stored, "inverse_entries": ninv, "ms": round(t_solve64, 3)}

    # -- the same FP64 step at the reference's default amalgamation (32, 0.25, 8): what a user of
    # the reference gets without tuning it (more, narrower supernodes and tree levels)
    if not args.no_default_amalgamation:
        sd = make_solver(csr, n, perm, (32, 0.25, 8), tiles=args.tiles)
        t_def = time_steps(sd, X, B, max(3, args.steps // 4), 1, 0)
        res["default_amalgamation"] = {"amalgamation": [32, 0.25, 8], "ms_per_step": round(t_def, 3),
                                       "backward_error": backward_error(A, X.cpu().numpy(), Bh),
                                       "supernodes": sd.symbolic.ns, "nnz_stored": sd.symbolic.nnz_stored,
                                       "autotuned_tiles": sd.tiles}
        del sd

    # -- FP32 factor + FP64 refinement
    f32_factor = None
    if not args.no_fp32:
        s32 = make_solver(csr, n, perm, amal, "Float32", tiles=args.tiles)
        steps_needed = None
        for k in range(1, 11):
            sds.setparam(s32, "ir_n_steps", k)
            sds.execute("solve", s32, X, B)
            if backward_error(A, X.cpu().numpy(), Bh) <= 1e-10:
                steps_needed = k
                break
        if steps_needed is not None:
            t_step32 = time_steps(s32, X, B, args.steps, args.warmup, steps_needed)
            berr32 = backward_error(A, X.cpu().numpy(), Bh)
            t_refact32 = time_phase(s32, args.steps, lambda: sds.execute("refactorization", s32))
            t_graph32 = time_graph(s32, X, B, args.steps, args.warmup)
            res["graph"].update({"fp32_ir_ms_per_step": round(t_graph32, 3),
                                 "fp32_ir_backward_error": backward_error(A, X.cpu().numpy(), Bh),
                                 "fp32_ir_speedup_vs_fp64": round(t_graph64 / t_graph32, 3)})
            res["fp32_ir"] = {"ms_per_step": round(t_step32, 3), "ir_steps": steps_needed,
                              "refactor_ms": round(t_refact32, 3),
                              "factor_tflops": round(flops_true / (t_refact32 * 1e-3) / 1e12, 3),
                              "backward_error": berr32, "speedup_vs_fp64": round(t_step64 / t_step32, 3)}
        else:
            res["fp32_ir"] = {"ms_per_step": None, "note": "backward error 1e-10 not reached in 10 refinement steps"}
        f32_factor = numeric_factor(s32).astype(np.float64)
        del s32

    # -- CPU baseline (rank 0, N = 1) and the GPU factors' errors against its FP64 factor
    if not args.no_cpu:
        res["cpu_baseline"], fac_cpu = cpu_baseline(S, csr[2], A, Bh)
        scale = np.abs(fac_cpu).max()
        res["fp64"]["factor_rel_err_vs_cpu"] = float(np.abs(numeric_factor(s64) - fac_cpu).max() / scale)
        if f32_factor is not None:
            res["fp32_ir"]["factor_rel_err_vs_cpu_fp64"] = float(np.abs(f32_factor - fac_cpu).max() / scale)
        res["cpu_baseline"]["gpu_speedup"] = round(res["cpu_baseline"]["value"] / t_step64, 1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
