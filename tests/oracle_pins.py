"""Pins of the oracle's pivot choice against independent references.

Each `pin_*` function raises AssertionError when the oracle loaded by `oracle.lib()` does not
reproduce what an independent reference fixes. `test_oracle_pivots.py` runs them against the
real oracle, and again against deliberately broken builds of oracle.c (in a child process,
with SDS_ORACLE_LIB set) to show that each plausible transcription slip is caught.

The independent references:

* With one front (a dense matrix in natural order is one supernode), pivot threshold u = 0 and
  a tiny epsilon, the reference's Bunch-Kaufman rule (src/reference/ldlt.jl:280-325: lambda =
  the largest |a_ik| below the diagonal, 1x1 if |a_kk| >= alpha lambda, else sigma = the largest
  off-diagonal of column r, 1x1 if |a_kk| sigma >= alpha lambda^2, 1x1 at r if |a_rr| >= alpha
  sigma, else the 2x2 on (k, r), with the interchanges k <-> c and k+1 <-> r) is LAPACK's
  dsytf2 with uplo = 'L'. scipy's dsytrf with its default workspace (lwork = n) runs dsytf2.
* With one front and u = 1 the LU row choice (src/reference/lu.jl:216-231: keep the diagonal if
  |a_kk| >= u max_i |a_ik|, else the first fully summed row with the largest |a_ik|) is
  partial pivoting, LAPACK's dgetrf.
* On a front with contribution-block rows (f > w) the fully summed w x w block is only ever
  updated by itself, so with u = 0 (LDL^T) or u = 1 (LU) the pivots are those of dsytrf /
  dgetrf on the leading w x w block, whatever the rows below hold. Rows below that are 10x
  larger make any search that strays past row w pick a different pivot.
* Hand-derived known answers (the docstrings work them out) for pivot type 'D', for the
  best-1x1 fallback of Bunch-Kaufman with u > 0 whose threshold test reads the
  contribution-block rows, and for the LU threshold test with u < 1.
"""
import numpy as np
import scipy.linalg.lapack as lapack
import scipy.sparse as sp

import oracle
from sparsedirectsolver import matrices as M
from sparsedirectsolver import symbolic_analysis

K_1X1, K_2X2_FIRST, K_2X2_SECOND, K_PERTURBED = 1, 2, 3, 4
TINY = 1e-300


def _one_front(A, structure):
    """Analysis of a dense matrix in natural order: one supernode holding every column."""
    T = sp.csr_matrix(np.tril(A)) if structure != "G" else sp.csr_matrix(A)
    T.sort_indices()
    rp, cv, nz = M.csr_arrays(T)
    S = symbolic_analysis(rp, cv, A.shape[0], structure, "L" if structure != "G" else "F", ordering="natural")
    assert S.ns == 1 and S.front_ncols[0] == A.shape[0] and np.array_equal(S.perm, np.arange(1, A.shape[0] + 1))
    return S, nz


def sytf2_reference(A):
    """LAPACK's Bunch-Kaufman LDL^T of the symmetric A (lower): the local pivot order lp
    (0-based, after the interchanges), the pivot kinds, D's diagonal and 2x2 subdiagonal, and L
    with every interchange applied to the earlier columns too (the reference swaps whole rows
    of the front, LAPACK only the trailing block)."""
    n = A.shape[0]
    ldu, ipiv, info = lapack.dsytrf(np.asfortranarray(A), lower=1)
    assert info == 0
    lp = np.arange(n)
    kind = np.zeros(n, dtype=np.int8)
    d = np.zeros(n)
    e = np.zeros(n)
    L = np.eye(n)
    k = 0
    while k < n:
        if ipiv[k] > 0:
            p, q, step = k, ipiv[k] - 1, 1
            kind[k] = K_1X1
        else:
            assert ipiv[k + 1] == ipiv[k]
            p, q, step = k + 1, -ipiv[k] - 1, 2
            kind[k], kind[k + 1] = K_2X2_FIRST, K_2X2_SECOND
        if p != q:
            lp[[p, q]] = lp[[q, p]]
            L[[p, q], :k] = L[[q, p], :k]
        d[k] = ldu[k, k]
        if step == 2:
            d[k + 1] = ldu[k + 1, k + 1]
            e[k] = ldu[k + 1, k]
        L[k + step:, k:k + step] = ldu[k + step:, k:k + step]
        k += step
    return lp, kind, d, e, L


def getrf_reference(A):
    """LAPACK's partial-pivoting LU: the local row order (0-based), D = diag(U), the unit L and
    the unit U = D^-1 U_lapack."""
    n = A.shape[0]
    lu, piv, info = lapack.dgetrf(np.asfortranarray(A))
    assert info == 0
    lp = np.arange(n)
    for k in range(n):
        if piv[k] != k:
            lp[[k, piv[k]]] = lp[[piv[k], k]]
    d = np.diag(lu).copy()
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu, 1) / d[:, None] + np.eye(n)
    return lp, d, L, U


def _random_symmetric(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return (A + A.T) / 2


def _random_general(n, seed):
    return np.random.default_rng(seed).standard_normal((n, n))


BK_SIZES = (7, 20, 37, 60)
SEEDS = range(5)


def pin_bk_single_front():
    """Multifrontal oracle LDL^T (one front, u = 0, tiny epsilon) == dsytf2 on random symmetric
    indefinite matrices: pivot order and kinds exactly, D and L to rounding."""
    n2x2 = 0
    for n in BK_SIZES:
        for seed in SEEDS:
            A = _random_symmetric(n, 100 * n + seed)
            S, nz = _one_front(A, "S")
            F = oracle.ldlt(S, nz, pivot_type="B", pivot_threshold=0.0, eps=TINY)
            lp, kind, d, e, L = sytf2_reference(A)
            case = f"n={n} seed={seed}"
            assert np.array_equal(F["piv"], lp + 1), f"{case}: pivot order {F['piv']} != LAPACK {lp + 1}"
            assert np.array_equal(F["kind"], kind), f"{case}: pivot kinds differ from LAPACK"
            scale = max(np.abs(d).max(), np.abs(e).max())
            assert np.abs(F["d"][:n] - d).max() <= 1e-10 * scale, f"{case}: D differs from LAPACK"
            assert np.abs(F["d"][n:] - e).max() <= 1e-10 * scale, f"{case}: 2x2 subdiagonals differ"
            Lo = oracle.extract_L(S, F["factor"])
            assert np.abs(Lo - L).max() <= 1e-9 * max(1.0, np.abs(L).max()), f"{case}: L differs from LAPACK"
            n2x2 += int((kind == K_2X2_FIRST).sum())
    assert n2x2 > 20, "the random cases should exercise 2x2 pivots"


def pin_bk_cb_rows():
    """Dense oracle front with f > w (u = 0): the pivots are dsytf2's on the leading w x w block;
    the contribution-block rows (10x larger) must not enter the lambda / sigma searches."""
    for f, w in ((12, 7), (30, 18), (45, 40)):
        for seed in SEEDS:
            F0 = _random_symmetric(f, 7 * f + seed)
            F0[w:, :] *= 10.0
            F0[:, w:] *= 10.0
            R = oracle.ldlt_front(F0, w, pivot_type="B", pivot_threshold=0.0, eps=TINY)
            lp, kind, d, e, _ = sytf2_reference(F0[:w, :w])
            case = f"f={f} w={w} seed={seed}"
            assert np.array_equal(R["lp"], lp), f"{case}: pivot order {R['lp']} != LAPACK {lp}"
            assert np.array_equal(R["kind"], kind), f"{case}: pivot kinds differ from LAPACK"
            assert np.abs(R["d"] - d).max() <= 1e-10 * np.abs(d).max(), f"{case}: D differs from LAPACK"


def pin_lu_single_front():
    """Multifrontal oracle LU (one front, u = 1) == dgetrf: row order exactly, D, L, U to rounding."""
    for n in BK_SIZES:
        for seed in SEEDS:
            A = _random_general(n, 31 * n + seed)
            S, nz = _one_front(A, "G")
            F = oracle.lu(S, nz, pivoting=True, pivot_threshold=1.0, eps=TINY)
            lp, d, L, U = getrf_reference(A)
            case = f"n={n} seed={seed}"
            assert np.array_equal(F["piv"], lp + 1), f"{case}: row order differs from LAPACK"
            assert np.abs(F["d"][:n] - d).max() <= 1e-10 * np.abs(d).max(), f"{case}: D differs from LAPACK"
            Lo = oracle.extract_L(S, F["factor"])
            Uo = oracle.extract_L(S, F["ufactor"]).T
            assert np.abs(Lo - L).max() <= 1e-10, f"{case}: L differs from LAPACK"
            assert np.abs(Uo - U).max() <= 1e-9 * np.abs(U).max(), f"{case}: U differs from LAPACK"


def pin_lu_cb_rows():
    """Dense oracle LU front with f > w (u = 1): the row order is dgetrf's on the w x w block."""
    for f, w in ((12, 7), (30, 18)):
        for seed in SEEDS:
            F0 = _random_general(f, 11 * f + seed)
            F0[w:, :] *= 10.0
            R = oracle.lu_front(F0, w, pivoting=True, pivot_threshold=1.0, eps=TINY)
            lp, d, _, _ = getrf_reference(F0[:w, :w])
            case = f"f={f} w={w} seed={seed}"
            assert np.array_equal(R["lp"], lp), f"{case}: row order {R['lp']} != LAPACK {lp}"
            assert np.abs(R["d"] - d).max() <= 1e-10 * np.abs(d).max(), f"{case}: D differs from LAPACK"


KA_DIAGONAL = np.array([[1.0, 4, 0, 0, 0],
                        [4, 3, 1, 0, 0],
                        [0, 1, 6, 0, 0],
                        [0, 0, 0, 20, 50],
                        [0, 0, 0, 50, 7]])


def pin_known_answer_diagonal():
    """Pivot type 'D', u = 0.5, on KA_DIAGONAL (one front, f = w = 5).

    k = 0: |a00| = 1 < 0.5 * 4: rejected. Best acceptable 1x1 over columns 0..4: column 1
    (3 >= 0.5 * 4) then column 2 (6 > 3 and 6 >= 0.5 * 1); columns 3 and 4 are larger (20, 7) but
    fail the threshold (< 0.5 * 50). So c = 2, swap 0 <-> 2, d = 6, a11 <- 3 - 1/6 = 17/6.
    k = 1: 17/6 >= 0.5 * 4: kept. a22 <- 1 - (24/17) 4 = -79/17, kept at k = 2 (nothing below).
    k = 3: 20 < 25 and 7 < 25: no acceptable 1x1, the largest diagonal (20 at column 3) is taken.
    d = 20, a44 <- 7 - 2.5 * 50 = -118. Pivot order (1-based) [3, 2, 1, 4, 5], all 1x1,
    D = [6, 17/6, -79/17, 20, -118] (det of the blocks: -79 and -2360), inertia (3, 2)."""
    R = oracle.ldlt_front(KA_DIAGONAL, 5, pivot_type="D", pivot_threshold=0.5, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [3, 2, 1, 4, 5]), f"'D' pivot order {R['lp'] + 1}"
    assert np.array_equal(R["kind"], [K_1X1] * 5)
    assert np.allclose(R["d"], [6, 17 / 6, -79 / 17, 20, -118], rtol=1e-14, atol=0)
    assert tuple(R["stats"][:2]) == (3, 2), f"inertia {tuple(R['stats'][:2])}"


KA_FALLBACK = np.array([[2.0, 1, 10],
                        [1, 8, 3],
                        [10, 3, 1]])


def pin_known_answer_bk_fallback():
    """Bunch-Kaufman, u = 0.5, on KA_FALLBACK with w = 2 (row 2 is a contribution-block row).

    k = 0: lambda = |a10| = 1 (the block's rows only), |a00| = 2 >= alpha: the BK choice is the
    1x1 at column 0. Its threshold test reads the whole column: 2 < 0.5 * |a20| = 5, rejected.
    Best 1x1 fallback: column 0 fails, column 1 passes (8 >= 0.5 * max(1, 3)): c = 1, swap
    0 <-> 1, d = 8, l = [1/8, 3/8], a11 <- 2 - 1/8 = 15/8, a21 <- 10 - 3/8 = 77/8.
    k = 1: 15/8 < 0.5 * 77/8 fails the threshold and nothing else is left: the 1x1 is taken as
    is (it is not tiny). Pivot order [2, 1], kinds [1x1, 1x1], D = [8, 15/8], L21 = [3/8, 77/15].
    A search past row w (lambda over the contribution rows) or a threshold test over the block
    rows only gives another order."""
    R = oracle.ldlt_front(KA_FALLBACK, 2, pivot_type="B", pivot_threshold=0.5, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [2, 1]), f"fallback pivot order {R['lp'] + 1}"
    assert np.array_equal(R["kind"], [K_1X1, K_1X1])
    assert np.allclose(R["d"], [8, 15 / 8], rtol=1e-15, atol=0)
    assert np.allclose(R["F"][2, :2], [3 / 8, 77 / 15], rtol=1e-14, atol=0)


KA_LU_CB = np.array([[2.0, 1, 1],
                     [3, 4, 1],
                     [10, 1, 1]])
KA_LU_THRESHOLD = np.array([[4.0, 1],
                            [6, 5]])
KA_LU_TIE = np.array([[1.0, 2, 3],
                      [3, 1, 1],
                      [-3, 1, 2]])


def pin_known_answer_lu_threshold():
    """LU threshold test, u = 0.5.

    KA_LU_CB with w = 2 (row 2 is a contribution-block row): k = 0: gamma over all rows = 10,
    |a00| = 2 < 5, so the largest fully summed row is taken: row 1 (|3| > 2). Swap, d = 3,
    a11 <- 1 - (2/3) 4 = -5/3, a21 <- 1 - (10/3) 4 = -37/3, a12 <- 1 - (2/3) 1 = 1/3.
    k = 1: no fully summed row left below: d = -5/3, l21 = 37/5, u12 = -1/5.
    Row order [2, 1], D = [3, -5/3]. (gamma over the fully summed rows only would keep row 0.)

    KA_LU_THRESHOLD, one front: |a00| = 4 >= 0.5 * 6 keeps the diagonal (order [1, 2], D = [4,
    3.5]); with u = 1 the larger row 1 is taken: order [2, 1], d1 = 6, d2 = 1 - (4/6) 5 = -7/3.

    KA_LU_TIE, one front, u = 1: rows 1 and 2 tie at |3| in column 0 and the first is taken.
    Then a11 = 2 - 1/3 = 5/3 loses to a21 = 1 + 1 = 2: order [2, 3, 1], D = [3, 2, 1/6]
    (a22 = 8/3 - (5/6) 3; the product of D is det = 1, the order an even permutation)."""
    R = oracle.lu_front(KA_LU_CB, 2, pivoting=True, pivot_threshold=0.5, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [2, 1]), f"LU order {R['lp'] + 1}"
    assert np.allclose(R["d"], [3, -5 / 3], rtol=1e-15, atol=0)
    assert np.allclose([R["F"][1, 0], R["F"][2, 0], R["F"][2, 1]], [2 / 3, 10 / 3, 37 / 5], rtol=1e-14)
    assert np.allclose([R["F"][0, 1], R["F"][0, 2], R["F"][1, 2]], [4 / 3, 1 / 3, -1 / 5], rtol=1e-14)
    R = oracle.lu_front(KA_LU_THRESHOLD, 2, pivoting=True, pivot_threshold=0.5, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [1, 2]) and np.allclose(R["d"], [4, 3.5], rtol=1e-15)
    R = oracle.lu_front(KA_LU_THRESHOLD, 2, pivoting=True, pivot_threshold=1.0, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [2, 1]) and np.allclose(R["d"], [6, -7 / 3], rtol=1e-15)
    R = oracle.lu_front(KA_LU_TIE, 3, pivoting=True, pivot_threshold=1.0, eps=1e-13)
    assert np.array_equal(R["lp"] + 1, [2, 3, 1]), f"LU tie order {R['lp'] + 1}"
    assert np.allclose(R["d"], [3, 2, 1 / 6], rtol=1e-14)


PINS = {name: fn for name, fn in globals().items() if name.startswith("pin_")}


if __name__ == "__main__":
    # run every pin against the oracle that SDS_ORACLE_LIB names; print the ones that fail
    import sys
    failed = []
    for name, fn in PINS.items():
        try:
            fn()
        except AssertionError as exc:
            failed.append(name)
            print(f"PIN-FAILED {name}: {str(exc)[:200]}")
    print("PINS-DONE", ",".join(failed))
    sys.exit(0)
