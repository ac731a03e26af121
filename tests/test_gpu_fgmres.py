"""FGMRES refinement (ir_mode = "fgmres") through the C ABI of libsds_hip.

The reference's semantics (src/solve/refinement.jl:554-620, src/solver.jl:385-392): FGMRES on
the correction system, right-preconditioned by the factorization, at most ir_n_steps
iterations, an early return when every column's relative residual is already below ir_tol,
and the iteration count reported as the "ir_n_steps" data parameter. The reference runs it
through Krylov.jl, which is absent here; what pins it are properties that hold for any
correct GMRES: the residual it leaves is no larger than that of plain refinement with the
same number of preconditioner applications (both corrections lie in the same Krylov space,
and GMRES minimises over it), the stopping rule, the transposed LU solve, and bitwise
repeatability.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import sparsedirectsolver as sds
from generators import random_unsym
from sparsedirectsolver import matrices as M
from sparsedirectsolver.errors import InvalidValueError, NotSupportedError
from sparsedirectsolver.solver import graph_capture

pytestmark = pytest.mark.gpu


def _spd_solver(A, perm, fp="Float32", amalgamation=(512, 0.25, 8)):
    rp, cv, nz = M.csr_arrays(M.lower_csr(A))
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "SPD", "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    sds.setparam(s, "amalgamation", amalgamation)
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _berr(A, X, B):
    """Column-wise backward error ||B - A X||_inf / (||A||_inf ||X||_inf + ||B||_inf)."""
    R = B - A @ X
    return (np.abs(R).max(axis=0) / (abs(A).sum(axis=1).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max()


def _solve(s, B, mode, steps, tol=0.0):
    import torch
    sds.setparam(s, "ir_mode", mode)
    sds.setparam(s, "ir_n_steps", steps)
    sds.setparam(s, "ir_tol", tol)
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    return X.cpu().numpy(), sds.getparam(s, "ir_n_steps")


def _rhs(n, k, seed):
    import torch
    return torch.tensor(np.random.default_rng(seed).standard_normal((n, k)), dtype=torch.float64, device="cuda")


def test_fgmres_fp32_factor_reaches_fp64_accuracy():
    A = M.laplacian3d(16, 16, 16)
    s = _spd_solver(A, sds.grid_nd_perm(16, 16, 16))
    B = _rhs(A.shape[0], 8, 2)
    X, its = _solve(s, B, "fgmres", 6)
    assert its == 6  # ir_tol = 0: every iteration runs
    berr = _berr(A, X, B.cpu().numpy())
    print("fgmres FP32 factor, 6 iterations: backward error", berr)
    assert berr <= 1e-14


def test_fgmres_residual_not_above_plain_refinement():
    """A rough-coefficient Laplacian (conductivities exp(N(0, 2^2))) with an FP32 factor: after
    k preconditioner applications FGMRES's stacked residual is at most plain refinement's."""
    dims = (14, 13, 12)
    A = M.variable_laplacian(dims, seed=5, sigma=2.0)
    s = _spd_solver(A, sds.grid_nd_perm(*dims), amalgamation=(128, 0.25, 8))
    B = _rhs(A.shape[0], 4, 7)
    Bh = B.cpu().numpy()
    for k in (1, 2, 3):
        Xi, _ = _solve(s, B, "ir", k)
        Xg, its = _solve(s, B, "fgmres", k)
        ri = np.linalg.norm(Bh - A @ Xi)
        rg = np.linalg.norm(Bh - A @ Xg)
        print(f"k={k}: plain {ri:.3e} fgmres {rg:.3e} ({its} iterations)")
        assert rg <= 1.05 * ri + 1e-15 * np.linalg.norm(Bh)


def test_fgmres_stopping_rule():
    A = M.laplacian3d(12, 12, 12)
    s = _spd_solver(A, sds.grid_nd_perm(12, 12, 12))
    B = _rhs(A.shape[0], 3, 4)
    Bh = B.cpu().numpy()
    # every column already meets a loose tolerance after the sweeps: no iteration
    _, its = _solve(s, B, "fgmres", 10, tol=1e-2)
    assert its == 0
    # a tolerance reached before the iteration limit: early stop, every column below it
    tol = 1e-11
    X, its = _solve(s, B, "fgmres", 20, tol=tol)
    assert 0 < its < 20
    rel = np.linalg.norm(Bh - A @ X, axis=0) / np.linalg.norm(Bh, axis=0)
    assert rel.max() <= 2 * tol, rel


def test_fgmres_lu_both_solve_modes():
    """FP32 LU with threshold row interchanges, FGMRES for A x = b and A^T x = b."""
    from test_gpu_lu import _rng, _solver as _lu_solver
    n = 600
    A = (random_unsym(n, 0.01, _rng(3), 1.0) + 2.0 * sp.identity(n)).tocsr()
    s, _ = _lu_solver(A, None, (128, 0.5, 8), "A", 0.5, fp="Float32")
    sds.execute("factorization", s)
    B = _rhs(n, 2, 2)
    for mode in (0, 1):
        sds.setparam(s, "solve_mode", mode)
        X, its = _solve(s, B, "fgmres", 5)
        Am = A if mode == 0 else A.T.tocsr()
        berr = _berr(Am, X, B.cpu().numpy())
        assert its == 5
        assert berr <= 1e-14, (mode, berr)


def test_fgmres_repeatable_and_guarded():
    import torch
    A = M.laplacian3d(10, 10, 10)
    s = _spd_solver(A, sds.grid_nd_perm(10, 10, 10))
    B = _rhs(A.shape[0], 2, 9)
    X1, _ = _solve(s, B, "fgmres", 3)
    X2, _ = _solve(s, B, "fgmres", 3)
    assert np.array_equal(X1, X2)
    with pytest.raises(InvalidValueError):
        sds.setparam(s, "ir_mode", "gmres")
    # FGMRES synchronizes with the host every iteration: a graph cannot hold it
    sds.setparam(s, "ir_mode", "fgmres")
    sds.setparam(s, "ir_n_steps", 2)
    with pytest.raises(NotSupportedError):
        graph_capture(s, 2, "solve")
    sds.setparam(s, "ir_mode", "ir")
    torch.cuda.synchronize()


def test_fgmres_kkt_ldlt_fp32_factor():
    """The KKT generator of the benchmark ([H J^T; J -1e-8 I]) with an FP32 Bunch-Kaufman
    factor, whose tiny pivots are perturbed: FGMRES preconditioned by that factor reaches FP64
    accuracy, at least as well as plain refinement with the same number of solves."""
    import torch
    nh, nj = 1500, 750
    K = M.kkt_matrix(nh, nj)
    rp, cv, nz = M.csr_arrays(M.lower_csr(K))
    n = K.shape[0]
    s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
    sds.setparam(s, "amalgamation", (512, 0.25, 8))
    sds.setparam(s, "pivot_type", "A")
    sds.setparam(s, "pivot_threshold", 0.01)
    sds.setparam(s, "factor_precision", "Float32")
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    assert sds.getparam(s, "inertia") == (nh, nj)
    B = _rhs(n, 1, 5)
    Bh = B.cpu().numpy()
    Xg, its = _solve(s, B, "fgmres", 20, tol=1e-14)
    Xi, _ = _solve(s, B, "ir", its)
    bg, bi = _berr(K, Xg, Bh), _berr(K, Xi, Bh)
    print(f"KKT n={n} FP32 factor: fgmres {its} iterations berr {bg:.2e}, plain refinement {bi:.2e}")
    assert bg <= 1e-14
    assert bg <= 1.05 * bi + 1e-16


def test_early_exit_with_a_zero_rhs_column():
    """A zero right-hand-side column counts with its residual norm itself, and with 1 in the
    smallest ||B_k|| (_max_relative_residual and _min_rhs_norm, src/solve/refinement.jl:284,
    :544): both refinement modes stop early, and the zero column's solution stays zero."""
    import torch
    A = M.laplacian3d(14, 13, 12)
    s = _spd_solver(A, sds.grid_nd_perm(14, 13, 12))
    B = _rhs(A.shape[0], 3, 21)
    B[:, 1] = 0.0
    for mode in ("ir", "fgmres"):
        X, it = _solve(s, B, mode, 20, tol=1e-13)
        assert 0 < it < 20, f"{mode}: {it} iterations"
        assert np.all(X[:, 1] == 0.0)
        # the zero column's backward error is 0/0; its exact zero solution is checked above
        assert _berr(A, X[:, [0, 2]], B.cpu().numpy()[:, [0, 2]]) <= 1e-14
