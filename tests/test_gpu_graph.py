"""Refactorization + solve captured in one HIP graph (sds_graph_capture) and replayed: every
replay, with new values written in place and eager work and synchronizations between the
replays, gives bitwise the result of the eager refactorization + solve on the same values;
Cholesky (FP64 and FP32 + refinement), LDL^T and LU."""
import numpy as np
import pytest

import oracle
import sparsedirectsolver as sds
from generators import kkt_matrix
from sparsedirectsolver import matrices as M
from sparsedirectsolver.solver import graph_capture, graph_replay, numeric_factor, numeric_pivots

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind in ("spd", "spd32"):
        A = M.variable_laplacian((10, 9, 8), seed=2)
        L = M.lower_csr(A)
        rp, cv, nz = M.csr_arrays(L)
        s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "SPD", "L")
        sds.setparam(s, "user_perm", sds.grid_nd_perm(10, 9, 8))
        sds.setparam(s, "amalgamation", (128, 0.25, 8))
        if kind == "spd32":
            sds.setparam(s, "factor_precision", "Float32")
            sds.setparam(s, "ir_n_steps", 2)
    elif kind == "ldlt":
        A = kkt_matrix(200, 80, 1e-8, np.random.default_rng(3))
        L = M.lower_csr(A)
        rp, cv, nz = M.csr_arrays(L)
        s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "S", "L")
        sds.setparam(s, "amalgamation", (64, 0.4, 8))
    else:
        A = M.convection_diffusion3d(7, peclet=10.0).tocsr()
        A.sort_indices()
        rp, cv, nz = M.csr_arrays(A)
        s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "G", "F")
    sds.execute("analysis", s)
    return A, s, nz


def _oracle_check(kind, s, nz, Bh, Xg):
    """The replay's factor and solution against the CPU oracle on the same values (the solver's
    default pivoting parameters): the Cholesky factor to 100 eps, the LDL^T and LU factors, pivot
    order and kinds bitwise; the solution to 1e-10 relative, 1e-8 after the FP32 factor's
    refinement, and 1e-6 for the KKT system (delta = 1e-8 makes it ill-conditioned enough that
    two backward-stable solves with different summation orders differ at about 1e-9)."""
    S = s.symbolic
    if kind in ("spd", "spd32"):
        info, F, _ = oracle.cholesky(S, nz)
        assert info == 0
        if kind == "spd":
            assert np.abs(numeric_factor(s) - F).max() <= 100 * np.finfo(np.float64).eps * np.abs(F).max()
        Xr = oracle.solve_cholesky(S, F, Bh)
    else:
        F = oracle.ldlt(S, nz) if kind == "ldlt" else oracle.lu(S, nz)
        piv, pkind = numeric_pivots(s)
        assert np.array_equal(piv, F["piv"]) and np.array_equal(pkind, F["kind"])
        assert np.array_equal(numeric_factor(s), F["factor"])
        if kind == "lu":
            assert np.array_equal(numeric_factor(s, 1), F["ufactor"])
        Xr = oracle.solve_ldlt(S, F, Bh) if kind == "ldlt" else oracle.solve_lu(S, F, Bh)
    tol = {"spd": 1e-10, "spd32": 1e-8, "ldlt": 1e-6, "lu": 1e-10}[kind]
    assert np.abs(Xg - Xr).max() <= tol * np.abs(Xr).max()


@pytest.mark.parametrize("kind", ["spd", "spd32", "ldlt", "lu"])
def test_graph_replays_match_eager(kind):
    import torch
    A, s, nz = _case(kind)
    n = A.shape[0]
    rng = np.random.default_rng(5)
    nrhs = 3
    print(f"graph {kind} n={n}", flush=True)
    sds.execute("factorization", s)
    graph_capture(s, nrhs=nrhs)
    base = s.nzval.clone()
    for rep in range(3):
        scale = 1.0 + 0.25 * rep
        s.nzval.copy_(base * scale)          # new values, same buffer
        B = torch.tensor(rng.standard_normal((n, nrhs)), dtype=torch.float64, device="cuda")
        Xg = torch.zeros_like(B)
        info = graph_replay(s, Xg, B, check_info=True)
        assert info == 0
        torch.cuda.synchronize()
        # the replay's factor and solution against the oracle on the same values
        _oracle_check(kind, s, (base * scale).cpu().numpy(), B.cpu().numpy(), Xg.cpu().numpy())
        # the eager path on the same values
        Xe = torch.zeros_like(B)
        sds.execute("refactorization", s)
        sds.execute("solve", s, Xe, B)
        torch.cuda.synchronize()
        assert torch.equal(Xg, Xe), f"replay {rep}: max diff {(Xg - Xe).abs().max().item():.3e}"


def test_graph_ldlt_with_pivot_sign():
    """LDL^T with requested perturbation signs (pivot_sign): the signs go to the device before the
    capture begins, and every replay gives bitwise the eager refactorization + solve. A large
    pivot_epsilon perturbs many pivots, so the requested signs decide the inertia."""
    import torch
    A = kkt_matrix(200, 80, 1e-8, np.random.default_rng(3))
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    n = A.shape[0]
    s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
    sds.setparam(s, "amalgamation", (64, 0.4, 8))
    sds.setparam(s, "pivot_epsilon", 50.0)
    sds.execute("analysis", s)
    inertia = {}
    for sign in (1, -1):
        sds.setparam(s, "pivot_sign", np.full(n, sign, dtype=np.int8))
        sds.execute("factorization", s)
        inertia[sign] = sds.getparam(s, "inertia")
    assert sds.getparam(s, "npivots") > 0 and inertia[1] != inertia[-1]
    sds.setparam(s, "pivot_sign", np.where(np.arange(n) % 3 == 0, -1, 1).astype(np.int8))
    rng = np.random.default_rng(9)
    sds.execute("factorization", s)
    graph_capture(s, nrhs=2)
    for rep in range(3):
        B = torch.tensor(rng.standard_normal((n, 2)), dtype=torch.float64, device="cuda")
        Xg = torch.zeros_like(B)
        assert graph_replay(s, Xg, B, check_info=True) == 0
        torch.cuda.synchronize()
        Xe = torch.zeros_like(B)
        sds.execute("refactorization", s)
        sds.execute("solve", s, Xe, B)
        torch.cuda.synchronize()
        assert torch.equal(Xg, Xe), f"replay {rep}: max diff {(Xg - Xe).abs().max().item():.3e}"
