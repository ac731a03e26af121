"""Adjoint of x = A^-1 b (the rrule of the solver): the gradients of the loss 0.5 ||x - x*||^2 with
respect to b and to the stored values of A, from the transposed solve on the GPU factor and the
HIP scatter of -lambda x^T, checked against central finite differences on 64 random stored
entries and against the gradient computed with the CPU oracle's solves; FP64 and FP32 + IR.
"""
import numpy as np
import pytest

import oracle
import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M
from sparsedirectsolver.adjoint import SparseSolve, rrule_solve

pytestmark = pytest.mark.gpu


def _setup(kind, fp=None):
    if kind == "spd":
        A = M.variable_laplacian((9, 8, 7), seed=5)
        L = M.lower_csr(A)
        rp, cv, nz = M.csr_arrays(L)
        s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "SPD", "L")
        rows = np.repeat(np.arange(L.shape[0]), np.diff(L.indptr))
        cols = L.indices
    else:
        A = M.convection_diffusion3d(8, peclet=10.0).tocsr()
        A.sort_indices()
        rp, cv, nz = M.csr_arrays(A)
        s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "G", "F")
        rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
        cols = A.indices
    if fp:
        sds.setparam(s, "factor_precision", fp)
        sds.setparam(s, "ir_n_steps", 8)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return A, s, nz, rows, cols


def _loss_and_x(s, nz, b, xstar):
    import torch
    sds.update(s, torch.tensor(nz, dtype=torch.float64, device="cuda"))
    sds.execute("refactorization", s)
    x = torch.zeros_like(b)
    sds.execute("solve", s, x, b)
    x = x.cpu().numpy()
    return 0.5 * float(np.sum((x - xstar) ** 2)), x


@pytest.mark.parametrize("kind", ["spd", "lu"])
def test_adjoint_matches_finite_differences_and_oracle(kind):
    import torch
    A, s, nz, rows, cols = _setup(kind)
    n = A.shape[0]
    rng = np.random.default_rng(7)
    bh, xstar = rng.standard_normal(n), rng.standard_normal(n)
    b = torch.tensor(bh, dtype=torch.float64, device="cuda")
    x, pullback = rrule_solve(s, b)
    xbar = x - torch.tensor(xstar, dtype=torch.float64, device="cuda")
    bbar, abar = pullback(xbar)
    bbar, abar = bbar.cpu().numpy(), abar.cpu().numpy()
    xh = x.cpu().numpy()
    # gradient with the oracle's factor and solves
    S = s.symbolic
    xb = xh - xstar
    if kind == "spd":
        _, F, _ = oracle.cholesky(S, nz)
        lam = oracle.solve_cholesky(S, F, xb)
        g = -(lam[rows] * xh[cols] + np.where(rows != cols, lam[cols] * xh[rows], 0.0))
    else:
        F = oracle.lu(S, nz, pivot_threshold=0.01, eps=1e-13)
        lam = oracle.solve_lu(S, F, xb, transpose=True)
        g = -lam[rows] * xh[cols]
    assert np.abs(bbar - lam).max() <= 1e-10 * np.abs(lam).max()
    assert np.abs(abar - g).max() <= 1e-10 * np.abs(g).max()
    # central finite differences on 64 random stored entries, and on 8 entries of b
    picks = rng.choice(nz.shape[0], 64, replace=False)
    for p in picks:
        h = 1e-6 * max(1.0, abs(nz[p]))
        zp, zm = nz.copy(), nz.copy()
        zp[p] += h
        zm[p] -= h
        fd = (_loss_and_x(s, zp, b, xstar)[0] - _loss_and_x(s, zm, b, xstar)[0]) / (2 * h)
        assert abs(fd - abar[p]) <= 1e-5 * max(np.abs(abar).max(), 1e-12), (p, fd, abar[p])
    _loss_and_x(s, nz, b, xstar)
    for i in rng.choice(n, 8, replace=False):
        h = 1e-6
        bp, bm = bh.copy(), bh.copy()
        bp[i] += h
        bm[i] -= h
        fd = (_loss_and_x(s, nz, torch.tensor(bp, device="cuda"), xstar)[0]
              - _loss_and_x(s, nz, torch.tensor(bm, device="cuda"), xstar)[0]) / (2 * h)
        assert abs(fd - bbar[i]) <= 1e-5 * np.abs(bbar).max()


@pytest.mark.parametrize("kind", ["spd", "lu"])
def test_adjoint_fp32_refinement_and_autograd(kind):
    """FP32 factor + FP64 refinement: the gradients agree with the FP64 ones; the
    torch.autograd.Function gives the same gradients as the rrule."""
    import torch
    A, s64, nz, _, _ = _setup(kind)
    _, s32, _, _, _ = _setup(kind, fp="Float32")
    n = A.shape[0]
    rng = np.random.default_rng(9)
    b = torch.tensor(rng.standard_normal(n), dtype=torch.float64, device="cuda")
    xstar = torch.tensor(rng.standard_normal(n), dtype=torch.float64, device="cuda")
    sds.setparam(s64, "ir_n_steps", 2)
    x64, pb64 = rrule_solve(s64, b)
    g64b, g64a = pb64(x64 - xstar)
    x32, pb32 = rrule_solve(s32, b)
    g32b, g32a = pb32(x32 - xstar)
    assert (g32a - g64a).abs().max().item() <= 1e-8 * g64a.abs().max().item()
    assert (g32b - g64b).abs().max().item() <= 1e-8 * g64b.abs().max().item()
    nzt = torch.tensor(nz, dtype=torch.float64, device="cuda", requires_grad=True)
    bt = b.clone().requires_grad_(True)
    x = SparseSolve.apply(s64, nzt, bt)
    loss = 0.5 * ((x - xstar) ** 2).sum()
    loss.backward()
    assert (nzt.grad - g64a).abs().max().item() <= 1e-12 * g64a.abs().max().item()
    assert (bt.grad - g64b).abs().max().item() <= 1e-12 * g64b.abs().max().item()
