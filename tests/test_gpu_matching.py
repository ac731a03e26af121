"""Matching and scaling (src/matching/mc64.jl, src/solver.jl:543-604) and 2x2 pivot pairs
(src/symbolic/pairs.jl) through the GPU numeric phase (C ABI of libsds_hip).

With a matching the factorization factors the scaled values r_i a_ij c_j of the matched matrix
(the analysis of M = (Dr A Dc)[:, q] for "G", D A D for the symmetric structures): its pivots,
kinds, D and panels must be bitwise the CPU oracle's factorization of those values on the same
analysis (w .* nz is the same IEEE product on both sides). The solves answer the original system
A x = b (and A^T x = b for "G"), checked by the backward error with the unscaled matrix; Sylvester's
law fixes the inertia of D A D. Pivot pairs change only the analysis; the oracle runs on it.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sparsedirectsolver as sds
from generators import kkt_matrix, random_unsym
from sparsedirectsolver import matrices as M
from sparsedirectsolver.analysis import entry_scaling
from sparsedirectsolver.solver import numeric_diag, numeric_factor, numeric_pivots, numeric_stats

pytestmark = pytest.mark.gpu


def _rng(s):
    return np.random.default_rng(s)


def _berr(A, X, B):
    an = abs(A).sum(axis=1).max()
    return float(np.max(np.abs(B - A @ X).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))))


def _rhs(n, k, seed):
    import torch
    return torch.tensor(_rng(seed).standard_normal((n, k)), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("job", ["algo5", "algo4", "algo1"])
def test_symmetric_matching_kkt(job):
    """KKT [H J^T; J -1e-8 I] badly scaled by rows and columns of H: LDL^T of D A D (job 5) or of
    A (jobs 1-4 keep no scaling), bitwise against the oracle on the same values and analysis,
    the inertia (nh, nj), and the solution of A x = b to a backward error of 1e-12."""
    import torch
    nh, nj = 300, 120
    K = kkt_matrix(nh, nj, 1e-8, _rng(21))
    s_ = np.ones(nh + nj)
    s_[:nh] = np.exp(2 * _rng(22).standard_normal(nh))
    K = (sp.diags(s_) @ K @ sp.diags(s_)).tocsr()
    L = M.lower_csr(K)
    rp, cv, nz = M.csr_arrays(L)
    n = K.shape[0]
    s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
    sds.setparam(s, "matching_alg", job)
    sds.setparam(s, "amalgamation", (64, 0.25, 8))
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    print(f"job {job}: {s.symbolic.pivot_pairs.shape[0]} pairs", flush=True)
    m = s.matching
    assert m is not None and m.symmetric
    w = entry_scaling(m, rp, cv, n) if job == "algo5" else np.ones(len(nz))
    ref = oracle.ldlt(s.symbolic, w * nz, pivot_type="A", pivot_threshold=0.01, eps=1e-13)
    piv, kind = numeric_pivots(s)
    assert np.array_equal(piv, ref["piv"]) and np.array_equal(kind, ref["kind"])
    assert np.array_equal(numeric_diag(s), ref["d"]) and np.array_equal(numeric_factor(s), ref["factor"])
    assert sds.getparam(s, "inertia") == (nh, nj)
    assert np.array_equal(sds.getparam(s, "scale_row"), m.rscale)
    B = _rhs(n, 2, 5)
    X = torch.zeros_like(B)
    sds.setparam(s, "ir_n_steps", 3)
    sds.execute("solve", s, X, B)
    berr = _berr(K, X.cpu().numpy(), B.cpu().numpy())
    assert berr <= 1e-12, berr


@pytest.mark.parametrize("fp", [None, "Float32"])
def test_unsymmetric_matching_lu_both_solve_modes(fp):
    """LU of M = (Dr A Dc)[:, q] with the MC64 job-5 matching: an unsymmetric matrix with a zero
    diagonal and rows scaled over 1e+-6; FP64 bitwise against the oracle's LU of the scaled values
    on the matched analysis; A x = b and A^T x = b to a backward error of 1e-12 (FP32 factor with
    refinement: 1e-10)."""
    import torch
    n = 200
    A = random_unsym(n, 0.04, _rng(31), diag=0.0)
    A = (A + sp.diags(_rng(32).uniform(1, 2, n - 1), 1)).tocsr()  # keep it structurally nonsingular
    A = (sp.diags(np.exp(3 * _rng(33).standard_normal(n))) @ A).tocsr()
    A.sort_indices()
    rp, cv, nz = M.csr_arrays(A)
    s = sds.DirectSolver((rp, cv, nz, n), "G", "F")
    sds.setparam(s, "matching_alg", "algo5")
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    m = s.matching
    assert not m.symmetric and m.matched.all()
    if fp is None:
        w = entry_scaling(m, rp, cv, n)
        ref = oracle.lu(s.symbolic, w * nz, pivoting=True, pivot_threshold=0.01, eps=1e-13)
        piv, kind = numeric_pivots(s)
        assert np.array_equal(piv, ref["piv"])
        assert np.array_equal(numeric_factor(s), ref["factor"]) and np.array_equal(numeric_factor(s, 1), ref["ufactor"])
    sds.setparam(s, "ir_n_steps", 2 if fp is None else 6)
    for mode in (0, 1):
        sds.setparam(s, "solve_mode", mode)
        B = _rhs(n, 2, 7 + mode)
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        Am = A if mode == 0 else A.T.tocsr()
        berr = _berr(Am, X.cpu().numpy(), B.cpu().numpy())
        assert berr <= (1e-12 if fp is None else 1e-10), (mode, berr)


def test_pivot_pairs_kkt_delta0():
    """KKT with delta = 0 (a structurally zero (2,2) block): the pivot-pair analysis ("default")
    against none: LDL^T bitwise against the oracle on each analysis, the inertia (nh, nj), and
    fewer perturbed pivots with the pairs (their zero pivots taken as 2x2 pivots)."""
    import torch
    nh, nj = 400, 200
    K = kkt_matrix(nh, nj, 0.0, _rng(41))
    L = M.lower_csr(K)
    rp, cv, nz = M.csr_arrays(L)
    n = K.shape[0]
    npert = {}
    for mode in ("none", "default", "all"):
        s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
        sds.setparam(s, "pivot_pairs", mode)
        sds.setparam(s, "amalgamation", (64, 0.25, 8))
        sds.execute("analysis", s)
        sds.execute("factorization", s)
        ref = oracle.ldlt(s.symbolic, nz, pivot_type="A", pivot_threshold=0.01, eps=1e-13)
        piv, kind = numeric_pivots(s)
        assert np.array_equal(piv, ref["piv"]) and np.array_equal(kind, ref["kind"])
        assert np.array_equal(numeric_factor(s), ref["factor"])
        assert np.array_equal(numeric_stats(s)[0], ref["stats"])
        assert sds.getparam(s, "inertia") == (nh, nj)
        npert[mode] = int(numeric_stats(s)[1][3])
        print(f"pivot_pairs {mode}: {s.symbolic.pivot_pairs.shape[0]} pairs, {npert[mode]} perturbed pivots", flush=True)
        B = _rhs(n, 1, 9)
        X = torch.zeros_like(B)
        sds.setparam(s, "ir_n_steps", 4)
        sds.execute("solve", s, X, B)
        assert _berr(K, X.cpu().numpy(), B.cpu().numpy()) <= 1e-10
    assert npert["default"] <= npert["none"]
