"""oracle/reference_path.py (the reference's CPU path with LAPACK potrf / BLAS trsm, syrk and gemm
per front, the CPU baseline of the benchmark) against the C oracle, whose element loops it
replaces by library calls: the same factor to rounding, the same solution, the same failed
column (test/test_reference_cholesky.jl:121-124)."""
import numpy as np
import scipy.sparse as sp

import oracle
import reference_path
from generators import random_spd
from sparsedirectsolver import grid_nd_perm, symbolic_analysis
from sparsedirectsolver import matrices as M


def _analyze(A, perm=None, amalgamation=(32, 0.25, 8)):
    T = sp.tril(A, format="csr")
    T.sort_indices()
    rp, cv, nz = M.csr_arrays(T)
    return symbolic_analysis(rp, cv, A.shape[0], "SPD", "L", perm=perm, amalgamation=amalgamation), nz


def test_lapack_path_equals_the_oracle():
    rng = np.random.default_rng(4)
    for A, perm, amal in ((M.laplacian3d(12, 11, 10), grid_nd_perm(12, 11, 10), (512, 0.25, 8)),
                          (random_spd(300, 0.02, rng), None, (32, 0.25, 8)),
                          (M.laplacian2d(40, 30), None, (1, 0.0, 0))):
        S, nz = _analyze(A, perm, amal)
        info, fac = reference_path.cholesky(S, nz)
        info_o, fac_o, _ = oracle.cholesky(S, nz)
        assert info == info_o == 0
        assert np.abs(fac - fac_o).max() <= 1e-13 * np.abs(fac_o).max()
        B = rng.standard_normal((A.shape[0], 3))
        X = reference_path.solve(S, fac, B)
        Xo = oracle.solve_cholesky(S, fac_o, B)
        assert np.abs(X - Xo).max() <= 1e-12 * np.abs(Xo).max()
        assert np.abs(A @ X - B).max() <= 1e-10 * np.abs(B).max()


def test_lapack_path_failed_pivot():
    A = sp.csr_matrix(np.array([[4, 0, 0, 0], [0, 4, 0, 0], [0, 0, 1, 2], [0, 0, 2, 1]], dtype=float))
    S, nz = _analyze(A, np.arange(1, 5))
    info, _ = reference_path.cholesky(S, nz)
    assert info == 4
