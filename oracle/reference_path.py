"""The reference's CPU path with its own dense calls -- TEST INFRASTRUCTURE ONLY (the checker
and the CPU baseline of bench.py; the product path never imports it).

`cholesky` restates ref_factorize! (src/reference/cholesky.jl:44-111) front by front in
postorder: assembly of A and extend-add of the children's contribution blocks (the C loops of
oracle.c, as the reference's own loops), then LAPACK potrf of F11, BLAS trsm of F21 and syrk of
F22 (cholesky.jl:87-101), here scipy's LAPACK/BLAS (OpenBLAS) in place of Julia's
LinearAlgebra (also OpenBLAS). `solve` restates ref_solve! (cholesky.jl:127-170): trsm on each
diagonal block and a gemm for the rows below, forward in postorder and backward in reverse.
The BLAS thread count is the caller's (threadpoolctl), as Julia's BLAS.set_num_threads.
"""
import ctypes

import numpy as np
import scipy.linalg.blas as blas
import scipy.linalg.lapack as lapack

import oracle


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def cholesky(S, nzval):
    """Returns (info, factor) in the panel layout of the analysis (info: 0 or the original
    column of the first failed pivot, 1-based, as ref_factorize!)."""
    lib = oracle.lib()
    o = oracle.OracleSym(S)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    factor = np.zeros(S.factor_len)
    f_all, w_all = S.front_nrows, S.front_ncols
    fptr, spar, sptr, perm = S.front_ptr, S.snparent, S.super_ptr, S.perm
    ctx = lib.oracle_front_begin(o.ptr)
    try:
        for s1 in S.snpost:
            s = int(s1) - 1
            f, w = int(f_all[s]), int(w_all[s])
            m = f - w
            F = np.empty((f, f), order="F")
            lib.oracle_front_assemble(ctx, s, _p(nz), _p(F))
            L11, info = lapack.dpotrf(F[:w, :w], lower=1, clean=1, overwrite_a=1)
            if info > 0:
                return int(perm[sptr[s] - 1 + info - 1]), factor
            panel = factor[fptr[s] - 1:fptr[s] - 1 + f * w].reshape((f, w), order="F")
            panel[:w] = L11
            if m > 0:
                # F21 <- F21 L11^-T ('R', 'L', 'C', 'N'), then C = -F21 F21^T into the parent's block
                L21 = blas.dtrsm(1.0, L11, F[w:, :w], side=1, lower=1, trans_a=1, diag=0)
                panel[w:] = L21
                if spar[s] != 0:
                    C = blas.dsyrk(-1.0, L21, lower=1)
                    lib.oracle_front_keep(ctx, s, _p(F), _p(np.asfortranarray(C)))
    finally:
        lib.oracle_front_end(ctx)
    return 0, factor


def solve(S, factor, B):
    """X = A^-1 B with a factor of `cholesky` (B: n or n x nrhs)."""
    B = np.asarray(B, dtype=np.float64)
    vec = B.ndim == 1
    B2 = B.reshape(S.n, -1)
    perm = S.perm - 1
    Y = np.asfortranarray(B2[perm])
    f_all, w_all, fptr, sptr = S.front_nrows, S.front_ncols, S.front_ptr, S.super_ptr
    rows_all, rptr = S.rowval - 1, S.rowptr - 1

    def panel(s):
        f, w = int(f_all[s]), int(w_all[s])
        P = factor[fptr[s] - 1:fptr[s] - 1 + f * w].reshape((f, w), order="F")
        return f, w, int(sptr[s]) - 1, rows_all[rptr[s] + w:rptr[s + 1]], P

    for s1 in S.snpost:                      # forward: L Z = Y
        f, w, c0, below, P = panel(int(s1) - 1)
        Yc = blas.dtrsm(1.0, P[:w], Y[c0:c0 + w], lower=1)
        Y[c0:c0 + w] = Yc
        if f > w:
            Y[below] -= P[w:] @ Yc
    for s1 in S.snpost[::-1]:                # backward: L^T W = Z
        f, w, c0, below, P = panel(int(s1) - 1)
        Yc = Y[c0:c0 + w]
        if f > w:
            Yc = Yc - P[w:].T @ Y[below]
        Y[c0:c0 + w] = blas.dtrsm(1.0, P[:w], Yc, lower=1, trans_a=1)
    X = np.empty_like(Y)
    X[perm] = Y
    return X.reshape(B.shape) if not vec else X[:, 0]
