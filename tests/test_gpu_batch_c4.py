"""The uneven-batch workload at its full size: 1,024 2-D 5-point variable-coefficient Laplacians,
sides log-uniform in [24, 140] (sum n = 5,783,419), conductivities exp(N(0, 1)), each with its own
nested dissection, all members in one block-diagonal forest (PLAN.md:635-639 of the reference).

- FP64: the batch factor equals the CPU oracle's factor of the same forest to 100 eps relative
  (every member), the per-front statistics are equal, and every member's backward error is at
  most 1e-14.
- FP32 factor + FP64 refinement, tile shapes from the autotuner, refactorization + solve replayed
  from one HIP graph: every replay equals the eager refactorization + solve bitwise, and every
  member's backward error is at most 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M
from sparsedirectsolver.batch import UnevenBatch
from sparsedirectsolver.solver import graph_capture, graph_replay, numeric_factor, numeric_stats

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
COUNT = 1024
AMALGAMATION = (128, 0.25, 8)


def _batch(fp=None):
    rng = np.random.default_rng(666)
    sides = np.round(np.exp(rng.uniform(np.log(24), np.log(140), COUNT))).astype(int)
    mats = [M.lower_csr(M.variable_laplacian((int(q), int(q)), seed=1000 + k)) for k, q in enumerate(sides)]
    perms = [sds.grid_nd_perm(int(q), int(q), 1) for q in sides]
    batch = UnevenBatch(mats, "SPD", "L", perms=perms)
    batch.setparam("amalgamation", AMALGAMATION)
    if fp:
        batch.setparam("factor_precision", fp)
    batch.execute("analysis")
    return batch


def _member_berr(batch, X, B):
    """Backward error ||b_k - A_k x_k||_inf / (||A_k||_inf ||x_k||_inf + ||b_k||_inf) of every
    member k (one right-hand side)."""
    A = batch.A + batch.A.T - sp.diags(batch.A.diagonal())
    off = batch.offsets[:-1]
    X, B = np.ravel(X), np.ravel(B)
    R = np.abs(B - A @ X)
    rows = np.asarray(abs(A).sum(axis=1)).ravel()
    return np.maximum.reduceat(R, off) / (np.maximum.reduceat(rows, off) * np.maximum.reduceat(np.abs(X), off) +
                                          np.maximum.reduceat(np.abs(B), off))


def test_c4_fp64_factor_matches_the_oracle():
    import torch
    batch = _batch()
    s = batch.solver
    print(f"C4 FP64: n={batch.n} supernodes={s.symbolic.ns}", flush=True)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    nz = s.nzval.cpu().numpy()
    info, ref, stats = oracle.cholesky(s.symbolic, nz)
    assert info == 0
    got = numeric_factor(s)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"C4 FP64: factor error vs the oracle {err:.2e}", flush=True)
    assert err <= 100 * EPS
    assert np.array_equal(numeric_stats(s)[0], stats)
    b = np.random.default_rng(7).standard_normal(batch.n)
    B = torch.tensor(b, dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    berr = _member_berr(batch, X.cpu().numpy(), b)
    assert berr.shape == (COUNT,)
    assert berr.max() <= 1e-14, berr.max()


def test_c4_fp32_ir_graph_replay_matches_eager():
    import torch
    batch = _batch("Float32")
    s = batch.solver
    sds.setparam(s, "ir_n_steps", 2)
    print(f"C4 FP32+IR: n={batch.n}, autotuning", flush=True)
    tiles = sds.autotune(s, reps=1)
    print(f"C4 FP32+IR: tiles {tiles}", flush=True)
    rng = np.random.default_rng(9)
    b = rng.standard_normal((batch.n, 1))
    B = torch.tensor(b, dtype=torch.float64, device="cuda")
    sds.execute("factorization", s)
    Xe = torch.zeros_like(B)
    sds.execute("solve", s, Xe, B)
    torch.cuda.synchronize()
    Xe = Xe.cpu().numpy()
    berr = _member_berr(batch, Xe, b)
    print(f"C4 FP32+IR: largest member backward error {berr.max():.2e}", flush=True)
    assert berr.max() <= 1e-10
    graph_capture(s, nrhs=1)
    for rep in range(3):
        Xg = torch.zeros_like(B)
        assert graph_replay(s, Xg, B, check_info=True) == 0
        torch.cuda.synchronize()
        assert np.array_equal(Xg.cpu().numpy(), Xe), f"replay {rep}"
        # eager work between the replays
        Xt = torch.zeros_like(B)
        sds.execute("solve", s, Xt, B)
        torch.cuda.synchronize()
