"""GPU LU with threshold partial pivoting (C ABI of libsds_hip) against the CPU oracle.

The device kernels perform the CPU reference's floating-point operations in the same order
(src/reference/lu.jl:84-231), so in FP64 the row orders, D, the L and U^T panels and the
statistics must be bitwise equal to the oracle's. Solves with A and A^T are checked by their
backward error against the oracle's solves, and the CUDSS.jl documentation example must give
x = ones (test/matrices.jl:331-345).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sparsedirectsolver as sds
from generators import random_unsym
from sparsedirectsolver import matrices as M
from sparsedirectsolver.solver import numeric_diag, numeric_factor, numeric_pivots, numeric_stats

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _solver(A, perm=None, amalgamation=(32, 0.25, 8), pivot_type="A", threshold=0.01, eps=None, fp=None):
    A = A.tocsr()
    A.sort_indices()
    rp, cv, nz = M.csr_arrays(A)
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "G", "F")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    sds.setparam(s, "amalgamation", amalgamation)
    sds.setparam(s, "pivot_type", pivot_type)
    sds.setparam(s, "pivot_threshold", threshold)
    if eps is not None:
        sds.setparam(s, "pivot_epsilon", eps)
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    return s, nz


def _berr(A, X, B):
    an = abs(A).sum(axis=1).max()
    return float(np.max(np.abs(B - A @ X).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))))


def _small_diag(A, step, value):
    A = A.tolil()
    for i in range(0, A.shape[0], step):
        A[i, i] = value
    return A.tocsr()


CASES = [
    # name, matrix, amalgamation, pivot_type, threshold
    ("convdiff_6_pivoting", lambda: _small_diag(M.convection_diffusion3d(6, peclet=10.0), 7, 1e-3),
     (32, 0.25, 8), "A", 0.5),
    ("convdiff_10_big", lambda: _small_diag(M.convection_diffusion3d(10, peclet=10.0), 5, 1e-2),
     (128, 0.5, 8), "A", 0.5),
    ("unsym_300_big", lambda: random_unsym(300, 0.03, _rng(3), 0.1), (64, 0.5, 8), "A", 0.1),
    ("convdiff_14_wide", lambda: _small_diag(M.convection_diffusion3d(14, peclet=10.0), 9, 1e-2),
     (256, 0.5, 8), "A", 0.5),
    ("unsym_200_nopivot", lambda: random_unsym(200, 0.03, _rng(4), 5.0), (32, 0.25, 8), "N", 0.01),
    ("unsym_150_one_front", lambda: random_unsym(150, 0.05, _rng(5), 0.5), (100000, 1.0, 100000), "B", 0.01),
]


@pytest.mark.parametrize("name,make,amal,ptype,thr", CASES, ids=[c[0] for c in CASES])
def test_lu_bitwise_oracle(name, make, amal, ptype, thr):
    import torch
    A = make()
    n = A.shape[0]
    print(f"case {name} n={n}", flush=True)
    s, nz = _solver(A, None, amal, ptype, thr)
    S = s.symbolic
    sds.execute("factorization", s)
    ref = oracle.lu(S, nz, pivoting=ptype not in ("N", "D"), pivot_threshold=thr, eps=1e-13)
    assert sds.getparam(s, "info") == ref["info"]
    piv, kind = numeric_pivots(s)
    assert np.array_equal(piv, ref["piv"]), "row order differs"
    assert np.array_equal(kind, ref["kind"]), "pivot kinds differ"
    assert np.array_equal(numeric_diag(s), ref["d"]), "D differs"
    assert np.array_equal(numeric_stats(s)[0], ref["stats"])
    got = numeric_factor(s, 0)
    assert np.array_equal(got, ref["factor"]), f"L differs by {np.abs(got - ref['factor']).max():.3e}"
    gotu = numeric_factor(s, 1)
    assert np.array_equal(gotu, ref["ufactor"]), f"U differs by {np.abs(gotu - ref['ufactor']).max():.3e}"
    sds.execute("refactorization", s)
    assert np.array_equal(numeric_factor(s, 0), got) and np.array_equal(numeric_factor(s, 1), gotu)
    B = torch.tensor(_rng(666).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    Bh = B.cpu().numpy()
    for mode in (0, 1):
        sds.setparam(s, "solve_mode", mode)
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        Am = A if mode == 0 else A.T.tocsr()
        Xr = oracle.solve_lu(S, ref, Bh, transpose=mode == 1)
        assert _berr(Am, X.cpu().numpy(), Bh) <= 10 * _berr(Am, Xr, Bh) + 1e-14, f"solve_mode {mode}"


def test_lu_cudss_doc_example_ones():
    import torch
    rows = np.array([1, 1, 1, 2, 2, 3, 3, 4, 4, 4, 5, 5]) - 1
    cols = np.array([1, 3, 5, 2, 4, 2, 3, 1, 4, 5, 4, 5]) - 1
    vals = np.array([4, 1, 2, 5, 3, 6, 8, 7, 9, 1, 1, 10], dtype=float)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(5, 5))
    s, _ = _solver(A)
    sds.execute("factorization", s)
    b = torch.tensor([7.0, 8.0, 14.0, 17.0, 11.0], dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    sds.execute("solve", s, x, b)
    assert np.allclose(x.cpu().numpy(), np.ones(5), rtol=0, atol=1e-13)


def test_lu_fp32_refinement_both_modes():
    """FP32 factor (row choices on FP32 values) + FP64 refinement on a moderately conditioned
    (cond ~ 5e3) unsymmetric matrix with row interchanges and large fronts."""
    import torch
    n = 600
    A = (random_unsym(n, 0.01, _rng(3), 1.0) + 2.0 * sp.identity(n)).tocsr()
    s, _ = _solver(A, None, (128, 0.5, 8), "A", 0.5, fp="Float32")
    sds.execute("factorization", s)
    sds.setparam(s, "ir_n_steps", 6)
    B = torch.tensor(_rng(2).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    for mode in (0, 1):
        sds.setparam(s, "solve_mode", mode)
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        Am = A if mode == 0 else A.T.tocsr()
        assert _berr(Am, X.cpu().numpy(), B.cpu().numpy()) <= 1e-10, f"solve_mode {mode}"
