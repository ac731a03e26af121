"""Pins of the CPU oracle (oracle/oracle.c) against what the mathematics fixes and against
the reference's own known answers. Each test names the plausible mistake it catches."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from generators import (eigen_inertia, extract_ldlt, extract_lu, kkt_interleaved_perm, kkt_matrix, random_spd,
                        random_symindef)
from sparsedirectsolver import grid_nd_perm, symbolic_analysis
from sparsedirectsolver import matrices as M

EPS = np.finfo(np.float64).eps


def analyze(A, structure="SPD", view="L", perm=None, amalgamation=(32, 0.25, 8), ordering="default"):
    T = sp.tril(A, format="csr") if view == "L" else A.tocsr()
    T.sort_indices()
    rp, cv, nz = M.csr_arrays(T)
    S = symbolic_analysis(rp, cv, A.shape[0], structure, view, perm=perm, amalgamation=amalgamation,
                          ordering=ordering)
    return S, nz


# --------------------------------------------------------------------------- Cholesky

CHOL_CASES = [
    ("lap2d_12x9_natural", lambda: (M.laplacian2d(12, 9), None, (32, 0.25, 8))),
    ("lap2d_20x20_nd", lambda: (M.laplacian2d(20, 20), grid_nd_perm(20, 20, 1), (32, 0.25, 8))),
    ("lap3d_8_nd_wide", lambda: (M.laplacian3d(8, 8, 8), grid_nd_perm(8, 8, 8), (128, 0.25, 8))),
    ("random_spd_amd", lambda: (random_spd(150, 0.03, np.random.default_rng(1)), None, (32, 0.25, 8))),
    ("varlap_amd_noamalg", lambda: (M.variable_laplacian((9, 8, 5), seed=2), None, (1, 0.0, 0))),
]


@pytest.mark.parametrize("name,make", CHOL_CASES, ids=[c[0] for c in CHOL_CASES])
def test_cholesky_equals_dense(name, make):
    """Every stored panel entry equals the dense Cholesky factor of P A P^T (catches a dropped
    extend-add, a wrong relind/amap index, a transposed panel, a missing duplicate sum)."""
    A, perm, amal = make()
    S, nz = analyze(A, perm=perm, amalgamation=amal)
    info, fac, st = oracle.cholesky(S, nz)
    assert info == 0
    p = S.perm - 1
    Lc = np.linalg.cholesky(A.toarray()[np.ix_(p, p)])
    Lf = oracle.extract_L(S, fac)
    assert np.abs(Lf - Lc).max() <= 1e-13 * np.abs(Lc).max()
    # stats: npos = w per front, info 0
    assert np.array_equal(st.reshape(-1, 6)[:, 0], S.front_ncols)
    # the solve against scipy
    b = np.random.default_rng(0).uniform(-1, 1, (A.shape[0], 2))
    x = oracle.solve_cholesky(S, fac, b)
    xr = spla.spsolve(A.tocsc(), b)
    assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()


def test_cholesky_tridiagonal_closed_form():
    """tridiag(-1, 2, -1) in natural order: d_1 = 2, d_k = 2 - 1/d_{k-1} = (k+1)/k and
    L[k, k] = sqrt(d_k), L[k+1, k] = -1/L[k, k] (catches a sign error, a wrong pivot update)."""
    n = 40
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    S, nz = analyze(A, ordering="algo5", amalgamation=(1, 0.0, 0))
    info, fac, _ = oracle.cholesky(S, nz)
    L = oracle.extract_L(S, fac)
    k = np.arange(1, n + 1)
    assert np.allclose(np.diag(L), np.sqrt((k + 1) / k), rtol=1e-14)
    assert np.allclose(np.diag(L, -1), -1 / np.sqrt((k[:-1] + 1) / k[:-1]), rtol=1e-14)


def test_cholesky_failed_pivot_known_answer():
    """[4 0 0 0; 0 4 0 0; 0 0 1 2; 0 0 2 1]: info == 4 (test/test_reference_cholesky.jl:121-124)."""
    A = sp.csr_matrix(np.array([[4, 0, 0, 0], [0, 4, 0, 0], [0, 0, 1, 2], [0, 0, 2, 1]], dtype=float))
    S, nz = analyze(A, ordering="algo5")
    info, _, st = oracle.cholesky(S, nz)
    assert info == 4


def test_cholesky_prefix_is_the_full_factor_on_its_subtrees():
    """The sample used for the CPU baseline computes the same panels as the full run."""
    A = M.laplacian3d(10, 9, 8)
    S, nz = analyze(A, perm=grid_nd_perm(10, 9, 8))
    _, full, _ = oracle.cholesky(S, nz)
    q1 = S.ns // 2
    _, part, _ = oracle.cholesky_prefix(S, nz, q1)
    done = S.snpost[:q1] - 1
    for s in done:
        a, b = S.front_ptr[s] - 1, S.front_ptr[s + 1] - 1
        assert np.array_equal(part[a:b], full[a:b])


# --------------------------------------------------------------------------- LDL^T

def _ldlt_check(A, S, F, u=0.01):
    L, D, p = extract_ldlt(S, F)
    Ad = A.toarray()[np.ix_(p, p)]
    err = np.linalg.norm(Ad - L @ D @ L.T) / np.linalg.norm(Ad)
    assert err <= 1e-10, err
    st = F["stats"].reshape(-1, 6).sum(axis=0)
    return st


LDLT_CASES = [
    ("symindef_300", lambda rng: (random_symindef(300, 0.02, rng), None, (32, 0.25, 8))),
    ("symindef_250_natural", lambda rng: (random_symindef(250, 0.05, rng), "natural", (32, 0.25, 8))),
    ("kkt_200_100_spd", lambda rng: (kkt_matrix(200, 100, 1e-2, rng), None, (32, 0.25, 8))),
    ("kkt_indef_interleaved", lambda rng: (kkt_matrix(200, 100, 0.0, rng, "indefinite"),
                                           kkt_interleaved_perm(200, 100), (32, 0.25, 8))),
    ("kkt_2x2_one_front", lambda rng: (kkt_matrix(120, 60, 0.0, rng, "indefinite", 1e-3), None,
                                       (10_000, 100.0, 10_000))),
    ("kkt_2x2_interleaved", lambda rng: (kkt_matrix(200, 100, 0.0, rng, "indefinite", 1e-3),
                                         kkt_interleaved_perm(200, 100), (32, 0.25, 8))),
]


@pytest.mark.parametrize("name,make", LDLT_CASES, ids=[c[0] for c in LDLT_CASES])
def test_ldlt_inertia_equals_eigenvalue_signs(name, make):
    """Inertia of D == eigenvalue signs of A (Sylvester), ||P A P^T - L D L^T|| small, solve
    accurate (test/test_reference_ldlt.jl:97-134). Catches a wrong 2x2 determinant sign rule, a
    dropped symmetric swap, a wrong update of the trailing block."""
    rng = np.random.default_rng(666)
    A, perm, amal = make(rng)
    ordering = "algo5" if isinstance(perm, str) else "default"
    S, nz = analyze(A, "S", perm=None if isinstance(perm, str) else perm, amalgamation=amal, ordering=ordering)
    F = oracle.ldlt(S, nz)
    assert F["info"] == 0
    st = _ldlt_check(A, S, F)
    npos, nneg, nzero = eigen_inertia(A)
    assert nzero == 0
    assert st[3] == 0
    assert (st[0], st[1]) == (npos, nneg)
    assert st[0] + st[1] == A.shape[0]
    if "2x2" in name:
        assert st[4] > 0
        n = A.shape[0]
        ks = np.where(F["kind"] == 2)[0]
        assert np.all(F["kind"][ks + 1] == 3) and np.all(F["d"][n + ks] != 0)
    b = rng.uniform(-1, 1, A.shape[0])
    x = oracle.solve_ldlt(S, F, b)
    assert np.linalg.norm(A @ x - b) / max(np.linalg.norm(b), 1) <= np.sqrt(EPS)


def test_ldlt_no_pivoting_on_spd_is_cholesky_squared():
    """pivot_type 'N' on an SPD matrix: D = diag(L_chol)^2 and L = L_chol diag(L_chol)^-1."""
    A = M.variable_laplacian((7, 6, 5), seed=4)
    S, nz = analyze(A, "S")
    F = oracle.ldlt(S, nz, pivot_type="N")
    _, fac, _ = oracle.cholesky(S, nz)
    Lc = oracle.extract_L(S, fac)
    n = A.shape[0]
    assert np.allclose(F["d"][:n], np.diag(Lc) ** 2, rtol=1e-12)
    assert np.array_equal(F["piv"], np.arange(1, n + 1))


def test_ldlt_threshold_bounds_l():
    """Every accepted 1x1 pivot satisfies the threshold test, so |L| <= 1/u entrywise for the
    1x1 columns (catches an inverted threshold comparison)."""
    rng = np.random.default_rng(3)
    A = random_symindef(200, 0.05, rng)
    S, nz = analyze(A, "S", amalgamation=(10_000, 100.0, 10_000))
    u = 0.1
    F = oracle.ldlt(S, nz, pivot_threshold=u)
    L, D, p = extract_ldlt(S, F)
    one = F["kind"] == 1
    assert np.abs(L[:, one] - np.eye(A.shape[0])[:, one]).max() <= 1 / u + 1e-12


def test_ldlt_zero_pivot_fails_with_zero_epsilon():
    """An exactly zero pivot with pivot_epsilon = 0 is reported as the failed column
    (test/test_numeric_ldlt.jl:111-143): [0 0; 0 1] natural, pivot 'N'."""
    A = sp.csr_matrix(np.array([[0.0, 0.0], [0.0, 1.0]]))
    A = A + sp.csr_matrix((np.array([0.0]), (np.array([0]), np.array([0]))), shape=(2, 2))
    T = sp.tril(A, format="csr")
    T = sp.csr_matrix((np.array([0.0, 1.0]), (np.array([0, 1]), np.array([0, 1]))), shape=(2, 2))
    rp, cv, nz = M.csr_arrays(T)
    S = symbolic_analysis(rp, cv, 2, "S", "L", ordering="algo5")
    F = oracle.ldlt(S, nz, pivot_type="N", eps=0.0)
    assert F["info"] == 1
    F2 = oracle.ldlt(S, nz, pivot_type="N", eps=1e-8)
    assert F2["info"] == 0 and F2["stats"].reshape(-1, 6)[:, 3].sum() == 1 and F2["d"][0] == 1e-8


# --------------------------------------------------------------------------- LU

def test_lu_reconstruction_and_solves():
    """A[p, q] = L D U and both A x = b and A^T x = b (catches a row swap applied to the wrong
    block, U^T stored untransposed, a missing local pivot order in the solve)."""
    rng = np.random.default_rng(5)
    A = M.convection_diffusion3d(6, peclet=10.0)
    # make some diagonal entries small so that the threshold test swaps rows
    A = A.tolil()
    for i in range(0, A.shape[0], 7):
        A[i, i] = 1e-3
    A = A.tocsr()
    S, nz = analyze(A, "G", view="F")
    F = oracle.lu(S, nz, pivot_threshold=0.5)
    assert F["info"] == 0
    L, d, U, p, q = extract_lu(S, F)
    Ad = A.toarray()[np.ix_(p, q)]
    # the tolerance of the reference's LDL^T/LU factor checks with element growth (test/utils.jl:574)
    assert np.linalg.norm(Ad - L @ np.diag(d) @ U) / np.linalg.norm(Ad) <= 1e-10
    assert np.any(F["piv"] != np.arange(1, A.shape[0] + 1))
    b = rng.standard_normal(A.shape[0])
    an = abs(A).sum(axis=1).max()

    def berr(M_, x):
        return np.abs(M_ @ x - b).max() / (an * np.abs(x).max() + np.abs(b).max())

    x = oracle.solve_lu(S, F, b)
    assert berr(A, x) <= 1e-14
    assert np.allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-8)
    xt = oracle.solve_lu(S, F, b, transpose=True)
    assert berr(A.T, xt) <= 1e-14
    assert np.allclose(xt, spla.spsolve(A.T.tocsc(), b), rtol=1e-8)


def test_lu_cudss_doc_example_known_answer():
    """The 5x5 general example of the CUDSS.jl docs: b = [7, 8, 14, 17, 11] has the solution
    x = ones (test/matrices.jl:331-345)."""
    rows = np.array([1, 1, 1, 2, 2, 3, 3, 4, 4, 4, 5, 5]) - 1
    cols = np.array([1, 3, 5, 2, 4, 2, 3, 1, 4, 5, 4, 5]) - 1
    vals = np.array([4, 1, 2, 5, 3, 6, 8, 7, 9, 1, 1, 10], dtype=float)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(5, 5))
    S, nz = analyze(A, "G", view="F")
    F = oracle.lu(S, nz)
    x = oracle.solve_lu(S, F, np.array([7, 8, 14, 17, 11], dtype=float))
    assert np.allclose(x, np.ones(5), rtol=0, atol=1e-13)


def test_ldlt_cudss_doc_example_known_answer():
    """The 5x5 symmetric example of the CUDSS.jl docs: b = [7, 5, 4, 6, 5], x = ones
    (test/matrices.jl:355-366)."""
    rows = np.array([1, 1, 1, 2, 2, 3, 3, 4, 4, 4, 5, 5, 5]) - 1
    cols = np.array([1, 3, 5, 2, 4, 1, 3, 2, 4, 5, 1, 4, 5]) - 1
    vals = np.array([4, 1, 2, 2, 3, 1, 3, 3, 2, 1, 2, 1, 2], dtype=float)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(5, 5))
    S, nz = analyze(A, "S", view="L")
    F = oracle.ldlt(S, nz)
    x = oracle.solve_ldlt(S, F, np.array([7, 5, 4, 6, 5], dtype=float))
    assert np.allclose(x, np.ones(5), rtol=0, atol=1e-13)


def test_cholesky_cudss_doc_example_known_answer():
    """The 5x5 SPD example of the CUDSS.jl docs (upper view): b = [4.5, 4.5, 3, 3, 2], x = ones
    (test/matrices.jl:370-381)."""
    rows = np.array([1, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5]) - 1
    cols = np.array([1, 2, 3, 1, 2, 4, 1, 3, 2, 4, 5]) - 1
    vals = np.array([2.5, 1, 1, 1, 2.5, 1, 1, 2, 1, 2, 2], dtype=float)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(5, 5))
    U = sp.triu(A, format="csr")
    U.sort_indices()
    rp, cv, nz = M.csr_arrays(U)
    S = symbolic_analysis(rp, cv, 5, "SPD", "U")
    info, fac, _ = oracle.cholesky(S, nz)
    x = oracle.solve_cholesky(S, fac, np.array([4.5, 4.5, 3, 3, 2]))
    assert info == 0 and np.allclose(x, np.ones(5), rtol=0, atol=1e-13)


def test_lu_ubatch_example():
    """The 3x3 uniform-batch example of the CUDSS.jl docs, A(l) = [1+l 0 3; 4 5+l 0; 2 6 2+l]
    for l in (1, 10, -20), b = 1:3 per member (test/matrices.jl:384-400), against scipy."""
    for lam in (1, 10, -20):
        A = sp.csr_matrix(np.array([[1 + lam, 0, 3], [4, 5 + lam, 0], [2, 6, 2 + lam]], dtype=float))
        S, nz = analyze(A, "G", view="F")
        F = oracle.lu(S, nz)
        b = np.array([1.0, 2.0, 3.0])
        x = oracle.solve_lu(S, F, b)
        assert np.allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-13, atol=1e-14)
