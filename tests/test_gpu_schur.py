"""Schur complement mode (schur_mode = 1 with user_schur_indices; src/numeric/schur.jl,
src/solver.jl:400-416) against dense linear algebra on small matrices, for Cholesky ("SPD"),
Bunch-Kaufman LDL^T ("S", a KKT matrix with delta = 1e-2: the test is of the Schur machinery, not of conditioning) and threshold-pivoting LU ("G", convection-diffusion):

- getparam(solver, "schur_matrix") equals S = A22 - A21 A11^-1 A12 of the flagged rows and columns
  in increasing order;
- "solve_fwd_schur" leaves the condensed right-hand side b2 - A21 A11^-1 b1 in the last ns entries
  (factor order); with the solution of S x2 = b2' put there, ("solve_diag" and) "solve_bwd_schur"
  give the solution of A x = b;
- LDL^T: the inertia is that of A11 (the Schur block's pivots count for nothing), checked against
  the eigenvalue signs of A11;
- a full "solve" is refused in Schur mode.
"""
import numpy as np
import pytest

import sparsedirectsolver as sds
from generators import kkt_matrix
from sparsedirectsolver import matrices as M
from sparsedirectsolver.errors import NotSupportedError

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == "spd":
        A = M.variable_laplacian((12, 11), seed=4).tocsr()
        return A, M.lower_csr(A), "SPD", "L"
    if kind == "ldlt":
        A = kkt_matrix(140, 60, 1e-2, np.random.default_rng(5)).tocsr()
        return A, M.lower_csr(A), "S", "L"
    A = M.convection_diffusion3d(6, peclet=10.0).tocsr()
    A.sort_indices()
    return A, A, "G", "F"


@pytest.mark.parametrize("fp", [None, "Float32"])
@pytest.mark.parametrize("kind", ["spd", "ldlt", "lu"])
def test_schur_complement_and_condensed_solve(kind, fp):
    import torch
    A, stored, structure, view = _case(kind)
    n = A.shape[0]
    Ad = A.toarray()
    rng = np.random.default_rng(11)
    flags = np.zeros(n, dtype=np.int8)
    flags[rng.choice(n, 23, replace=False)] = 1
    i2 = np.flatnonzero(flags)
    i1 = np.flatnonzero(flags == 0)
    A11, A12, A21, A22 = Ad[np.ix_(i1, i1)], Ad[np.ix_(i1, i2)], Ad[np.ix_(i2, i1)], Ad[np.ix_(i2, i2)]
    S_ref = A22 - A21 @ np.linalg.solve(A11, A12)
    rp, cv, nz = M.csr_arrays(stored)
    s = sds.DirectSolver((rp, cv, nz, n), structure, view)
    sds.setparam(s, "amalgamation", (32, 0.25, 8))
    sds.setparam(s, "schur_mode", 1)
    sds.setparam(s, "user_schur_indices", flags)
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    ns = len(i2)
    assert sds.getparam(s, "schur_shape")[:2] == (ns, ns)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    tol = 1e-11 if fp is None else 2e-5
    Sg = sds.getparam(s, "schur_matrix").cpu().numpy()
    err = np.abs(Sg - S_ref).max() / np.abs(S_ref).max()
    print(f"{kind} fp={fp} n={n} ns={ns}: |S - S_ref| / |S_ref| = {err:.2e}", flush=True)
    assert err <= tol
    if kind == "ldlt" and fp is None:
        ev = np.linalg.eigvalsh(A11)
        assert sds.getparam(s, "inertia") == (int((ev > 0).sum()), int((ev < 0).sum()))
    b = rng.standard_normal(n)
    B = torch.tensor(b, dtype=torch.float64, device="cuda")
    # the condensed right-hand side
    Y = torch.zeros_like(B)
    sds.execute("solve_fwd_schur", s, Y, B)
    bs = Y[-ns:].cpu().numpy()
    bs_ref = b[i2] - A21 @ np.linalg.solve(A11, b[i1])
    assert np.abs(bs - bs_ref).max() <= tol * max(np.abs(bs_ref).max(), 1.0)
    # the Schur system on the host, then the backward phases
    x2 = np.linalg.solve(Sg, bs)
    Y[-ns:] = torch.tensor(x2, dtype=torch.float64, device="cuda")
    W = torch.zeros_like(Y)
    sds.execute("solve_diag", s, W, Y)
    X = torch.zeros_like(B)
    sds.execute("solve_bwd_schur", s, X, W)
    x = X.cpu().numpy()
    x_ref = np.linalg.solve(Ad, b)
    xerr = np.abs(x - x_ref).max() / np.abs(x_ref).max()
    print(f"{kind} fp={fp}: solution error {xerr:.2e}", flush=True)
    assert xerr <= (1e-9 if fp is None else 1e-4)
    with pytest.raises(NotSupportedError):
        sds.execute("solve", s, X, B)
