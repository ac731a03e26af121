"""Uniform batches (src/solver.jl:135-150, src/numeric/batch.jl): nbatch matrices with one
pattern, values member after member, right-hand sides with nrhs columns per member.

Each member must come out as if it were solved alone: its solution is compared bit for bit
with a single-matrix solver on the same values (the members are the blocks of one
block-diagonal matrix, and every front's arithmetic is the same as in the member's own
factorization). Per-member data parameters follow the reference: "info", "inertia" and
"npivots" are vectors over the members; "ubatch_index" / "ubatch_mask" restrict a phase to
some members and leave the others' factors and solutions untouched.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M
from sparsedirectsolver.errors import InvalidValueError

pytestmark = pytest.mark.gpu


def _members(nb, dims=(9, 8, 7)):
    mats = [M.lower_csr(M.variable_laplacian(dims, seed=10 + k)) for k in range(nb)]
    for m in mats[1:]:
        assert np.array_equal(m.indptr, mats[0].indptr) and np.array_equal(m.indices, mats[0].indices)
    return mats


def _batch_solver(mats, structure="SPD", perm=None, matrix_values=False):
    import torch
    rp, cv, _ = M.csr_arrays(mats[0])
    vals = np.stack([m.data for m in mats], axis=1)  # nnz x nbatch
    nz = torch.tensor(vals if matrix_values else vals.T.reshape(-1), dtype=torch.float64, device="cuda")
    s = sds.DirectSolver((rp, cv, nz, mats[0].shape[0]), structure, "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    sds.setparam(s, "amalgamation", (64, 0.25, 8))
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _single(m, structure="SPD", perm=None):
    rp, cv, nz = M.csr_arrays(m)
    s = sds.DirectSolver((rp, cv, nz, m.shape[0]), structure, "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    sds.setparam(s, "amalgamation", (64, 0.25, 8))
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _rhs(n, cols, seed):
    import torch
    return torch.tensor(np.random.default_rng(seed).standard_normal((n, cols)), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("matrix_values", [False, True])
def test_members_bitwise_as_solved_alone(matrix_values):
    import torch
    nb, r = 3, 2
    mats = _members(nb)
    n = mats[0].shape[0]
    s = _batch_solver(mats, matrix_values=matrix_values)
    assert s.nbatch == nb
    B = _rhs(n, r * nb, 1)
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    for k in range(nb):
        q = _single(mats[k])
        xk = torch.zeros((n, r), dtype=torch.float64, device="cuda")
        sds.execute("solve", q, xk, B[:, k * r:(k + 1) * r].contiguous())
        assert torch.equal(X[:, k * r:(k + 1) * r], xk), k
        # and the CPU oracle's solution of the member
        info, F, _ = oracle.cholesky(q.symbolic, mats[k].data)
        assert info == 0
        Xr = oracle.solve_cholesky(q.symbolic, F, B[:, k * r:(k + 1) * r].cpu().numpy())
        assert np.abs(xk.cpu().numpy() - Xr).max() <= 1e-10 * np.abs(Xr).max(), k
    assert sds.getparam(s, "info") == [0] * nb
    assert sds.getparam(s, "inertia") == [(n, 0)] * nb


def test_rhs_layouts_agree():
    import torch
    nb, r = 3, 2
    mats = _members(nb)
    n = mats[0].shape[0]
    s = _batch_solver(mats)
    B = _rhs(n, r * nb, 2)
    X2 = torch.zeros_like(B)
    sds.execute("solve", s, X2, B)
    # n x nrhs x nbatch array and the strided vector of n nrhs nbatch entries
    B3 = B.reshape(n, nb, r).permute(0, 2, 1).contiguous()
    X3 = torch.zeros_like(B3)
    sds.execute("solve", s, X3, B3)
    assert torch.equal(X3.permute(0, 2, 1).reshape(n, nb * r), X2)
    B1 = B.t().contiguous().reshape(-1)
    X1 = torch.zeros_like(B1)
    sds.execute("solve", s, X1, B1)
    assert torch.equal(X1.reshape(nb * r, n).t(), X2)
    with pytest.raises(InvalidValueError):
        sds.execute("solve", s, torch.zeros((n, 4), dtype=torch.float64, device="cuda"),
                    torch.zeros((n, 4), dtype=torch.float64, device="cuda"))


def test_ubatch_index_and_mask_leave_other_members_untouched():
    import torch
    nb, r = 3, 1
    mats = _members(nb)
    n = mats[0].shape[0]
    s = _batch_solver(mats)
    B = _rhs(n, r * nb, 3)
    X0 = torch.zeros_like(B)
    sds.execute("solve", s, X0, B)
    # new values for every member, but only member 1 is refactored
    nnz = mats[0].nnz
    new = [m.copy() for m in mats]
    for m in new:
        m.data = m.data * 2.0
    with torch.no_grad():
        s.nzval.copy_(torch.tensor(np.concatenate([m.data for m in new]), dtype=torch.float64, device="cuda"))
    sds.setparam(s, "ubatch_index", 1)
    sds.execute("refactorization", s)
    sds.setparam(s, "ubatch_index", -1)
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    assert torch.equal(X[:, 0], X0[:, 0]) and torch.equal(X[:, 2], X0[:, 2])
    q = _single(new[1])
    x1 = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    sds.execute("solve", q, x1, B[:, 1:2].contiguous())
    assert torch.equal(X[:, 1:2], x1)
    # a masked solve writes the active members' columns only
    sds.setparam(s, "ubatch_mask", [1, 0, 1])
    Xm = torch.full_like(B, 7.0)
    sds.execute("solve", s, Xm, B)
    assert torch.equal(Xm[:, 0], X[:, 0]) and torch.equal(Xm[:, 2], X[:, 2])
    assert torch.all(Xm[:, 1] == 7.0)
    sds.setparam(s, "ubatch_mask", [0, 0, 0])
    with pytest.raises(InvalidValueError):
        sds.execute("solve", s, Xm, B)
    sds.setparam(s, "ubatch_mask", None)
    assert nnz == mats[0].nnz


def test_per_member_info_and_inertia():
    import torch
    # member 1 is the reference's known answer: first non-positive pivot at column 4
    # (test/test_reference_cholesky.jl:121-124); member 0 shares its pattern and is SPD
    a0 = np.array([[4, 0, 0, 0], [0, 4, 0, 0], [0, 0, 3, 1], [0, 0, 1, 3]], dtype=float)
    a1 = np.array([[4, 0, 0, 0], [0, 4, 0, 0], [0, 0, 1, 2], [0, 0, 2, 1]], dtype=float)
    mats = [M.lower_csr(sp.csr_matrix(a)) for a in (a0, a1)]
    s = _batch_solver(mats, perm=np.arange(1, 5))
    assert sds.getparam(s, "info") == [0, 4]
    # LDL^T of A and -A: inertia per member
    A = M.laplacian2d(6, 5)
    L = M.lower_csr(A)
    Ln = L.copy()
    Ln.data = -Ln.data
    t = _batch_solver([L, Ln], structure="S")
    n = A.shape[0]
    assert sds.getparam(t, "inertia") == [(n, 0), (0, n)]
    assert sds.getparam(t, "info") == [0, 0]
    with pytest.raises(InvalidValueError):
        sds.setparam(t, "ubatch_size", 3)
        sds.execute("analysis", t)
    torch.cuda.synchronize()
