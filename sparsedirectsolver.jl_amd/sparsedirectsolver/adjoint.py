"""Differentiation of x = A^-1 b through the solver: a ChainRules-style ``rrule`` and a
``torch.autograd.Function``.

The reference has no differentiation rule (no ChainRules in its Project.toml); this is the
library's addition on the same factor. For x = A^-1 b and a cotangent xbar:

    bbar = lambda = A^-T xbar            (one transposed solve, solve_mode = 1, same factor)
    Abar[p] = -lambda_i x_j              (for the stored entry p at (i, j); both (i, j) and
                                          (j, i) for a mirrored entry of a symmetric structure)

The transposed solve runs through ``sds_solve`` and the scatter of -lambda x^T onto the stored
pattern through ``sds_adjoint_scatter`` (a HIP kernel, one thread per stored entry). With an FP32
factor the transposed solve uses the same FP64 iterative refinement as the forward solve.
"""
import ctypes

from .solver import _check, execute, getparam, setparam


def _torch():
    import torch
    return torch


def _as_cols(T, n):
    """(column-major n x nrhs view as a contiguous (nrhs, n) tensor, nrhs)"""
    torch = _torch()
    if T.dim() == 1:
        return T.reshape(1, n).contiguous().to(torch.float64), 1
    return T.t().contiguous().to(torch.float64), T.shape[1]


def adjoint_scatter(solver, lam, X):
    """Abar on the stored pattern: -lambda x^T gathered per stored entry (device, Float64, nnz)."""
    torch = _torch()
    n = solver.n
    L, nrhs = _as_cols(lam, n)
    Xc, nx = _as_cols(X, n)
    if nx != nrhs:
        raise ValueError("lambda and X must have the same number of right-hand sides")
    grad = torch.empty(solver.colval.shape[0], dtype=torch.float64, device=solver.nzval.device)
    _check(solver._lib, solver._h, solver._lib.sds_adjoint_scatter(
        solver._h, ctypes.c_void_p(L.data_ptr()), n, ctypes.c_void_p(Xc.data_ptr()), n, nrhs,
        ctypes.c_void_p(grad.data_ptr())))
    return grad


def transposed_solve(solver, Xbar):
    """lambda = A^-T xbar on the current factor (solve_mode = 1, the solver's refinement settings)."""
    torch = _torch()
    lam = torch.zeros_like(Xbar, dtype=torch.float64)
    mode = getparam(solver, "solve_mode")
    setparam(solver, "solve_mode", 1)
    try:
        execute("solve", solver, lam, Xbar.to(torch.float64))
    finally:
        setparam(solver, "solve_mode", mode)
    return lam


def rrule_solve(solver, B):
    """(X, pullback) with X = A^-1 B on the solver's current factor and
    pullback(Xbar) -> (Bbar, Abar_nzval), as ChainRules' rrule(\\, A, b)."""
    torch = _torch()
    X = torch.zeros_like(B, dtype=torch.float64)
    execute("solve", solver, X, B)

    def pullback(Xbar):
        lam = transposed_solve(solver, Xbar)
        return lam, adjoint_scatter(solver, lam, X)

    return X, pullback


class SparseSolve:
    """torch.autograd.Function: x = solve(nzval, b) with gradients for nzval and b.

    ``SparseSolve.apply(solver, nzval, b)`` refactorizes the analysed solver with the values
    ``nzval`` (a Float64 cuda tensor in the order of the user's CSR) and solves; backward uses
    the same factor for the transposed solve and the pattern scatter."""

    _fn = None

    @classmethod
    def apply(cls, solver, nzval, b):
        return cls._function().apply(solver, nzval, b)

    @classmethod
    def _function(cls):
        if cls._fn is None:
            torch = _torch()

            class _Solve(torch.autograd.Function):
                @staticmethod
                def forward(ctx, solver, nzval, b):
                    from .solver import update
                    update(solver, nzval.detach())
                    execute("refactorization" if solver.stage >= 3 else "factorization", solver)
                    x = torch.zeros_like(b, dtype=torch.float64)
                    execute("solve", solver, x, b.detach())
                    ctx.solver = solver
                    ctx.save_for_backward(x)
                    return x

                @staticmethod
                def backward(ctx, xbar):
                    (x,) = ctx.saved_tensors
                    lam = transposed_solve(ctx.solver, xbar.contiguous())
                    return None, adjoint_scatter(ctx.solver, lam, x), lam

            cls._fn = _Solve
        return cls._fn
