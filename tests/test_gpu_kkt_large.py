"""The KKT generator at n = 60,000 ([H J^T; J -1e-8 I], H 40,000 x 40,000 with about 10 entries per
row, J 20,000 x 40,000 with 4 per row), LDL^T with Bunch-Kaufman pivoting, threshold 0.01: a size
where the CPU oracle would take hours, so the checks are the ones the mathematics fixes whatever
the size. By Sylvester's law the inertia is exactly (40000, 20000, 0): H is strictly diagonally
dominant with a positive diagonal (SPD), and with delta > 0 the Schur complement
-delta I - J H^-1 J^T is negative definite; the FP64 factor has that inertia. The FP64 solve with
refinement reaches a backward error of 1e-14; the FP32 factor with FGMRES refinement (the form
meant for these ill-conditioned systems) reaches the 1e-10 of the FP32 form's target (about 8e-12
in 20 iterations: at this size some 8,000 perturbed pivots weaken the preconditioner, where 9
iterations reach 1e-18 at n = 12,000). The FP32 factor's inertia is not checked: its
pivot_epsilon (1e-5) perturbs the tiny pivots of the -delta block to +-epsilon, which may change
their signs (the inertia counts the signs of D as factored)."""
import numpy as np
import pytest

import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M

pytestmark = pytest.mark.gpu


def _berr(A, x, b):
    return np.abs(b - A @ x).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("fp,mode", [("Float64", "ir"), ("Float32", "fgmres")])
def test_kkt_60000_inertia_and_backward_error(fp, mode):
    import torch
    nh, nj = 40000, 20000
    K = M.kkt_matrix(nh, nj)
    rp, cv, nz = M.csr_arrays(M.lower_csr(K))
    n = K.shape[0]
    s = sds.DirectSolver((rp, cv, nz, n), "S", "L")
    sds.setparam(s, "amalgamation", (512, 0.25, 8))
    sds.setparam(s, "pivot_type", "A")
    sds.setparam(s, "pivot_threshold", 0.01)
    if fp == "Float32":
        sds.setparam(s, "factor_precision", "Float32")
    sds.setparam(s, "ir_mode", mode)
    sds.setparam(s, "ir_n_steps", 3 if mode == "ir" else 20)
    if mode == "fgmres":
        sds.setparam(s, "ir_tol", 1e-14)
    sds.execute("analysis", s)
    print(f"KKT n={n} {fp}: nnz_stored {s.symbolic.nnz_stored}, flops {s.symbolic.flops_true:.3e}", flush=True)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    inertia = sds.getparam(s, "inertia")
    print(f"KKT n={n} {fp}: inertia {inertia}, perturbed pivots {sds.getparam(s, 'npivots')}", flush=True)
    if fp == "Float64":
        assert inertia == (nh, nj)
    b = np.random.default_rng(666).standard_normal(n)
    B = torch.tensor(b, dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    berr = _berr(K, X.cpu().numpy(), b)
    print(f"KKT n={n} {fp} {mode}: backward error {berr:.2e}, {s.ir_steps} refinement steps", flush=True)
    assert berr <= (1e-14 if fp == "Float64" else 1e-10)
