"""Uneven batches: the members of a block-diagonal forest are factored and solved by the same
level launches; each member's solution is bitwise the one of solving it alone (same ordering,
same fronts, same kernels), in FP64 and with the FP32 factor + FP64 refinement."""
import numpy as np
import pytest

import oracle
import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M
from sparsedirectsolver.batch import UnevenBatch
from sparsedirectsolver.solver import numeric_factor

pytestmark = pytest.mark.gpu


def _members(count, seed, lo=6, hi=40):
    rng = np.random.default_rng(seed)
    sides = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    mats = [M.lower_csr(M.variable_laplacian((int(s), int(s)), seed=int(seed * 1000 + k))) for k, s in enumerate(sides)]
    perms = [sds.grid_nd_perm(int(s), int(s), 1) for s in sides]
    return mats, perms


@pytest.mark.parametrize("fp", [None, "Float32"])
def test_batch_members_equal_individual_solves(fp):
    import torch
    mats, perms = _members(12, 3)
    print(f"members {[m.shape[0] for m in mats]} fp={fp}", flush=True)
    batch = UnevenBatch(mats, "SPD", "L", perms=perms)
    batch.setparam("amalgamation", (64, 0.25, 8))
    if fp:
        batch.setparam("factor_precision", fp)
        batch.setparam("ir_n_steps", 3)
    batch.execute("analysis")
    batch.execute("factorization")
    rng = np.random.default_rng(1)
    bs = [rng.standard_normal((m.shape[0], 2)) for m in mats]
    B = batch.stack(bs)
    X = torch.zeros_like(B)
    batch.execute("solve", X, B)
    for k, (m, p, b) in enumerate(zip(mats, perms, bs)):
        rp, cv, nz = M.csr_arrays(m)
        s = sds.DirectSolver((rp, cv, nz, m.shape[0]), "SPD", "L")
        sds.setparam(s, "user_perm", p)
        sds.setparam(s, "amalgamation", (64, 0.25, 8))
        if fp:
            sds.setparam(s, "factor_precision", fp)
            sds.setparam(s, "ir_n_steps", 3)
        sds.execute("analysis", s)
        sds.execute("factorization", s)
        bt = torch.tensor(b, dtype=torch.float64, device="cuda")
        x = torch.zeros_like(bt)
        sds.execute("solve", s, x, bt)
        assert torch.equal(batch.split(X)[k], x), f"member {k}"
    A = batch.A + batch.A.T - __import__("scipy").sparse.diags(batch.A.diagonal())
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    # the CPU oracle on the whole forest: its factor (FP64) and its solution
    info, F, _ = oracle.cholesky(batch.solver.symbolic, batch.solver.nzval.cpu().numpy())
    assert info == 0
    if fp is None:
        assert np.abs(numeric_factor(batch.solver) - F).max() <= 100 * np.finfo(np.float64).eps * np.abs(F).max()
    Xr = oracle.solve_cholesky(batch.solver.symbolic, F, Bh)
    assert np.abs(Xh - Xr).max() <= (1e-10 if fp is None else 1e-8) * np.abs(Xr).max()
    an = abs(A).sum(axis=1).max()
    berr = np.max(np.abs(Bh - A @ Xh).max(axis=0) / (an * np.abs(Xh).max(axis=0) + np.abs(Bh).max(axis=0)))
    assert berr <= (1e-14 if fp is None else 1e-10)
