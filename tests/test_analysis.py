"""The host analysis against brute force and the reference's known answers
(test/test_symbolic_supernodes.jl, test/test_symbolic_etree.jl)."""
import numpy as np
import pytest
import scipy.sparse as sp

from generators import random_spd, random_symindef
from sparsedirectsolver import grid_nd_perm, symbolic_analysis
from sparsedirectsolver import matrices as M


def _analyze(A, structure="SPD", perm=None, amalgamation=(32, 0.25, 8), use_superpanels=1, ordering="default"):
    T = sp.tril(A, format="csr")
    T.sort_indices()
    rp, cv, nz = M.csr_arrays(T)
    return symbolic_analysis(rp, cv, A.shape[0], structure, "L", perm=perm, amalgamation=amalgamation,
                             use_superpanels=use_superpanels, ordering=ordering)


def _symbolic_cholesky_pattern(A, perm):
    """Brute force: the structure of L of P A P^T by dense boolean elimination."""
    p = perm - 1
    B = (A.toarray()[np.ix_(p, p)] != 0)
    n = B.shape[0]
    for k in range(n):
        rows = np.where(B[k + 1:, k])[0] + k + 1
        for i in rows:
            B[rows, i] |= True
    return np.tril(B)


def test_arrow_matrix_known_answers():
    """test/test_symbolic_supernodes.jl:70-95: etree parent == [5, 5, 5, 5, 6, 0],
    supernodes [1, 2, 3, 4, 5:6], nnz_stored == nnz_L == 3 * 4 + 3; with amalgamation
    (8, 0.5, 8) one dense 6x6 supernode with nnz_stored == 21."""
    A = sp.csr_matrix(np.array([[4, 0, 0, 0, 1, 1], [0, 4, 0, 0, 1, 1], [0, 0, 4, 0, 1, 1], [0, 0, 0, 4, 1, 1],
                                [1, 1, 1, 1, 4, 1], [1, 1, 1, 1, 1, 4]], dtype=float))
    S = _analyze(A, perm=np.arange(1, 7), use_superpanels=0)
    assert list(S.parent) == [5, 5, 5, 5, 6, 0]
    assert list(S.super_ptr) == [1, 2, 3, 4, 5, 7]
    assert list(S.snparent) == [5, 5, 5, 5, 0]
    assert S.nnz_stored == S.nnz_L == 3 * 4 + 3
    assert list(S.snrows(1)) == [1, 5, 6] and list(S.snrows(5)) == [5, 6]
    Sa = _analyze(A, perm=np.arange(1, 7), amalgamation=(8, 0.5, 8))
    assert Sa.ns == 1 and Sa.nnz_stored == 21


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_structure_matches_brute_force(seed):
    """nnz_L (from the column counts) and every supernode's row set against a dense symbolic
    elimination (catches a wrong etree, colcount or supernode row union)."""
    rng = np.random.default_rng(seed)
    A = random_spd(60, 0.05, rng)
    for ordering in ("algo5", "default"):
        S = _analyze(A, ordering=ordering, use_superpanels=0)
        Lp = _symbolic_cholesky_pattern(A, S.perm)
        assert S.nnz_L == int(Lp.sum())
        assert np.array_equal(S.counts, Lp.sum(axis=0))
        for s in range(1, S.ns + 1):
            rows = S.snrows(s) - 1
            for j in S.sncols(s) - 1:
                true_rows = np.where(Lp[:, j])[0]
                assert set(true_rows) <= set(rows.tolist())
        # fundamental supernodes store no explicit zeros
        assert S.nnz_stored == S.nnz_L


def test_amalgamation_keeps_structure_superset_and_budget():
    A = M.laplacian3d(10, 10, 10)
    S0 = _analyze(A, perm=grid_nd_perm(10, 10, 10), use_superpanels=0)
    for amal in ((32, 0.25, 8), (512, 0.25, 8)):
        S = _analyze(A, perm=grid_nd_perm(10, 10, 10), amalgamation=amal)
        assert S.nnz_L == S0.nnz_L
        assert S.nnz_stored >= S.nnz_L
        assert S.nnz_stored - S.nnz_L <= amal[1] * S.nnz_L + 1
        Lp = _symbolic_cholesky_pattern(A, S.perm)
        for s in range(1, S.ns + 1):
            rows = set((S.snrows(s) - 1).tolist())
            for j in S.sncols(s) - 1:
                assert set(np.where(Lp[:, j])[0]) <= rows


def test_maps_place_every_entry():
    """amap + relind: assembling A through amap and summing children through relind gives back
    P A P^T's lower triangle (the symbolic half of the extend-add)."""
    rng = np.random.default_rng(7)
    A = random_symindef(80, 0.06, rng)
    S = _analyze(A, "S")
    T = sp.tril(A, format="csr")
    T.sort_indices()
    rp, cv, nz = M.csr_arrays(T)
    n = A.shape[0]
    F = np.zeros(S.factor_len)
    for k in range(S.amap_ptr[-1] - 1):
        p = S.amap_src[k] - 1
        F[abs(S.amap[p]) - 1] += nz[p]
    Ad = np.zeros((n, n))
    for s in range(1, S.ns + 1):
        rows = S.snrows(s) - 1
        f = rows.shape[0]
        c0 = S.super_ptr[s - 1] - 1
        base = S.front_ptr[s - 1] - 1
        for j in range(S.front_ncols[s - 1]):
            for i in range(j, f):
                Ad[rows[i], c0 + j] = F[base + j * f + i]
    p = S.perm - 1
    assert np.allclose(Ad, np.tril(A.toarray()[np.ix_(p, p)]))
    # relind: positions of the child's rows among its parent's
    for c in range(1, S.ns + 1):
        par = S.snparent[c - 1]
        if par == 0:
            continue
        rows_c = S.snrows(c)[S.front_ncols[c - 1]:]
        ri = S.relind[S.relind_ptr[c - 1] - 1:S.relind_ptr[c] - 1]
        assert np.array_equal(S.snrows(par)[ri - 1], rows_c)


def test_amd_reduces_fill():
    """The approximate minimum degree ordering is a permutation and beats the natural order on
    grids (quality check, not parity: AMD.jl is not available here)."""
    for A in (M.laplacian2d(30, 30), M.laplacian3d(10, 10, 10)):
        Sn = _analyze(A, ordering="algo5")
        Sa = _analyze(A, ordering="default")
        assert sorted(Sa.perm) == list(range(1, A.shape[0] + 1))
        assert Sa.nnz_L < Sn.nnz_L


def test_full_pattern_map_is_the_matrix():
    rng = np.random.default_rng(2)
    A = random_symindef(50, 0.1, rng)
    S = _analyze(A, "S")
    T = sp.tril(A, format="csr")
    T.sort_indices()
    nz = T.data
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(S.full_rowptr))
    vals = nz[np.abs(S.full_src) - 1]
    Mf = sp.csr_matrix((vals, (rows, S.full_colval - 1)), shape=(n, n)).toarray()
    assert np.allclose(Mf, A.toarray())
