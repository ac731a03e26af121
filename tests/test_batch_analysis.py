"""The block-diagonal forest of an uneven batch (analysis concatenation) on the CPU: with the
oracle, the batch factor is the concatenation of the members' factors, bit for bit, and so are
the statistics and the solutions; for Cholesky, LDL^T and LU maps."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from generators import random_symmetric, random_unsym
from sparsedirectsolver import matrices as M
from sparsedirectsolver.analysis import merge_symbolic, symbolic_analysis


def _analyse(m, structure, view):
    m = m.tocsr()
    m.sort_indices()
    rp, cv, nz = M.csr_arrays(m)
    return symbolic_analysis(rp, cv, m.shape[0], structure, view, "O", amalgamation=(64, 0.25, 8)), nz


@pytest.mark.parametrize("structure", ["SPD", "S", "G"])
def test_merged_forest_factor_is_the_members_factors(structure):
    rng = np.random.default_rng(4)
    if structure == "SPD":
        members = [M.lower_csr(M.variable_laplacian((s, s + 1), seed=s)) for s in (5, 9, 3, 14, 7)]
        view = "L"
    elif structure == "S":
        members = [M.lower_csr(random_symmetric(n, 0.08, rng, 0.05)) for n in (30, 55, 12, 80)]
        view = "L"
    else:
        members = [random_unsym(n, 0.08, rng, 0.3) for n in (25, 60, 9, 90)]
        view = "F"
    parts = [_analyse(m, structure, view) for m in members]
    S = merge_symbolic([p[0] for p in parts])
    nz = np.concatenate([p[1] for p in parts])
    B = sp.block_diag(members, format="csr")
    assert S.n == B.shape[0] and S.nnz == B.nnz
    b = rng.standard_normal(S.n)
    offs = np.concatenate([[0], np.cumsum([m.shape[0] for m in members])])
    if structure == "SPD":
        info, F, st = oracle.cholesky(S, nz)
        ref = [oracle.cholesky(p[0], p[1]) for p in parts]
        assert info == 0
        assert np.array_equal(F, np.concatenate([r[1] for r in ref]))
        assert np.array_equal(st, np.concatenate([r[2] for r in ref]))
        x = oracle.solve_cholesky(S, F, b)
        xs = [oracle.solve_cholesky(p[0], r[1], b[offs[k]:offs[k + 1]]) for k, (p, r) in enumerate(zip(parts, ref))]
    elif structure == "S":
        F = oracle.ldlt(S, nz)
        ref = [oracle.ldlt(p[0], p[1]) for p in parts]
        assert np.array_equal(F["factor"], np.concatenate([r["factor"] for r in ref]))
        assert np.array_equal(F["stats"], np.concatenate([r["stats"] for r in ref]))
        assert np.array_equal(F["piv"], np.concatenate([r["piv"] + o for r, o in zip(ref, offs[:-1])]))
        x = oracle.solve_ldlt(S, F, b)
        xs = [oracle.solve_ldlt(p[0], r, b[offs[k]:offs[k + 1]]) for k, (p, r) in enumerate(zip(parts, ref))]
    else:
        F = oracle.lu(S, nz)
        ref = [oracle.lu(p[0], p[1]) for p in parts]
        assert np.array_equal(F["factor"], np.concatenate([r["factor"] for r in ref]))
        assert np.array_equal(F["ufactor"], np.concatenate([r["ufactor"] for r in ref]))
        x = oracle.solve_lu(S, F, b, transpose=True)
        xs = [oracle.solve_lu(p[0], r, b[offs[k]:offs[k + 1]], transpose=True) for k, (p, r) in enumerate(zip(parts, ref))]
    assert np.array_equal(x, np.concatenate(xs))
