"""LDL^T and LU sweeps of large fronts with many right-hand sides.

The large-front sweeps (k_fwd_step, k_bwd_l21 + k_bwd_l21_sum, k_bwd_step) work on blocks of 8
right-hand sides and, in the backward product, on chunks of 256 rows below a front's columns with
the chunks' sums added in a fixed order. These cases have 11 right-hand sides (a full block of 8
and a ragged block of 3), fronts with more than 256 rows below their columns (several chunks) and
more than 64 columns (several block steps). No refinement runs, so the sweeps alone are checked:
every column's backward error against the matrix (1e-12 for the well-conditioned LU case; the
unrefined KKT factorization itself leaves about 1e-11), and the 11-column solve against the
solves of single columns (the same kernels over another block shape: the kernels accumulate each
right-hand side separately, so the results agree to rounding of the same operations).
"""
import numpy as np
import pytest

import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M

pytestmark = pytest.mark.gpu

NRHS = 11


def _berr_cols(A, X, B):
    an = abs(A).sum(axis=1).max()
    return np.abs(B - A @ X).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))


def _front_shape(S):
    """(largest rows below the columns, widest front) over the supernodes of the analysis."""
    below, width = 0, 0
    for k in range(1, S.ns + 1):
        f, w = len(S.snrows(k)), len(S.sncols(k))
        below, width = max(below, f - w), max(width, w)
    return below, width


def _check(s, A, modes, bound):
    import torch
    n = A.shape[0]
    below, width = _front_shape(s.symbolic)
    print(f"n={n}: {s.symbolic.ns} supernodes, at most {below} rows below a front, widest front {width}",
          flush=True)
    assert below > 256 and width > 64
    Bh = np.random.default_rng(12).standard_normal((n, NRHS))
    B = torch.tensor(Bh, dtype=torch.float64, device="cuda")
    for mode in modes:
        if mode is not None:
            sds.setparam(s, "solve_mode", mode)
        Am = A if not mode else A.T.tocsr()
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        Xh = X.cpu().numpy()
        berr = _berr_cols(Am, Xh, Bh)
        print(f"  solve_mode {mode}: largest backward error over the {NRHS} columns {berr.max():.2e}", flush=True)
        assert berr.max() <= bound
        for c in (0, 7, 8, NRHS - 1):
            b1 = B[:, c:c + 1].contiguous()
            x1 = torch.zeros_like(b1)
            sds.execute("solve", s, x1, b1)
            x1h = x1.cpu().numpy()[:, 0]
            rel = np.abs(x1h - Xh[:, c]).max() / np.abs(x1h).max()
            assert rel <= 1e-12, f"column {c}: {rel:.2e}"


def test_ldlt_sweeps_eleven_rhs():
    K = M.kkt_matrix(4000, 2000)
    rp, cv, nz = M.csr_arrays(M.lower_csr(K))
    n = K.shape[0]
    s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
    sds.setparam(s, "amalgamation", (512, 0.25, 8))
    sds.setparam(s, "reordering_alg", "algo3")
    sds.setparam(s, "pivot_type", "A")
    sds.setparam(s, "pivot_threshold", 0.01)
    sds.setparam(s, "ir_n_steps", 0)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    # unrefined LDL^T of this ill-conditioned system (delta = 1e-8, threshold 0.01: tiny pivots
    # of the -delta block) leaves a backward error of about 1e-11 in every column
    _check(s, K.tocsr(), (None,), 1e-10)


def test_lu_sweeps_eleven_rhs_both_modes():
    A = M.convection_diffusion3d(24, peclet=10.0).tocsr()
    A.sort_indices()
    rp, cv, nz = M.csr_arrays(A)
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "G", "F")
    sds.setparam(s, "amalgamation", (512, 0.25, 8))
    sds.setparam(s, "pivot_type", "A")
    sds.setparam(s, "pivot_threshold", 0.5)
    sds.setparam(s, "ir_n_steps", 0)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    _check(s, A, (0, 1), 1e-12)
