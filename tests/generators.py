"""Small test matrices following test/matrices.jl of the reference (random streams are numpy's,
so the matrices are of the same families, not the same draws)."""
import numpy as np
import scipy.sparse as sp


def sprand(rng, m, n, density):
    nnz = int(round(density * m * n))
    i = rng.integers(0, m, nnz)
    j = rng.integers(0, n, nnz)
    A = sp.coo_matrix((rng.uniform(0, 1, nnz), (i, j)), shape=(m, n)).tocsr()
    A.sum_duplicates()
    return A


def random_spd(n, density, rng):
    """B B^T + n I (test/matrices.jl:63-67)."""
    B = sprand(rng, n, n, density)
    A = (B @ B.T + n * sp.identity(n)).tocsr()
    return ((A + A.T) / 2).tocsr()


def random_symindef(n, density, rng):
    """off-diagonal B + B^T plus the diagonal +-(row-abs-sum + 1) with shuffled signs: by
    Gershgorin every |eigenvalue| >= 1 and the inertia is the sign count of the diagonal
    (test/matrices.jl:121-128)."""
    B = sprand(rng, n, n, density)
    S = (B + B.T).tolil()
    S.setdiag(0)
    S = S.tocsr()
    S.eliminate_zeros()
    r = np.asarray(abs(S).sum(axis=1)).ravel()
    signs = np.array([1 if i % 2 == 0 else -1 for i in range(n)])
    rng.shuffle(signs)
    return (S + sp.diags(signs * (r + 1))).tocsr()


def random_symmetric(n, density, rng, diag=0.0):
    """B + B^T with N(0, 1) entries and a small N(0, diag^2) diagonal (stored, possibly zero):
    strongly indefinite, so Bunch-Kaufman takes many 2x2 pivots."""
    B = sprand(rng, n, n, density)
    B.data = rng.standard_normal(B.data.size)
    S = (B + B.T).tocsr()
    S = S + sp.diags(diag * rng.standard_normal(n))
    S = (S + sp.csr_matrix((np.zeros(n), (np.arange(n), np.arange(n))), shape=(n, n))).tocsr()
    S.sort_indices()
    return S


def random_unsym(n, density, rng, diag=1.0):
    """Unsymmetric sparse B with N(0, 1) entries plus an N(0, diag^2) diagonal (stored)."""
    B = sprand(rng, n, n, density)
    B.data = rng.standard_normal(B.data.size)
    A = B + sp.diags(diag * rng.standard_normal(n))
    A = (A + sp.csr_matrix((np.zeros(n), (np.arange(n), np.arange(n))), shape=(n, n))).tocsr()
    A.sort_indices()
    return A


def kkt_matrix(nh, nj, delta, rng, hessian="spd", hessian_scale=1.0):
    """[H J^T; J -delta I] with J of full row rank (test/matrices.jl:143-158)."""
    H = random_spd(nh, min(1.0, 5 / nh), rng) if hessian == "spd" else random_symindef(nh, min(1.0, 5 / nh), rng)
    R = sprand(rng, nj, nh, min(1.0, 3 / nh))
    d = np.asarray(abs(R).sum(axis=1)).ravel() + 1
    J = (R + sp.diags(d, 0, shape=(nj, nh))).tocsr()
    D = sp.diags(np.full(nj, -delta))
    K = sp.bmat([[hessian_scale * H, J.T], [J, D]], format="csr")
    # the -delta I block is always stored, also for delta = 0
    K = K + sp.csr_matrix((np.zeros(nj), (np.arange(nh, nh + nj), np.arange(nh, nh + nj))), shape=K.shape)
    K.sort_indices()
    return K


def kkt_interleaved_perm(nh, nj):
    """primal i followed by its dual nh + i, then the remaining primals (1-based; test/matrices.jl:194)."""
    p = []
    for i in range(1, nj + 1):
        p += [i, nh + i]
    return np.array(p + list(range(nj + 1, nh + 1)), dtype=np.int64)


def eigen_inertia(A):
    ev = np.linalg.eigvalsh(A.toarray())
    tol = 1e-9 * np.abs(ev).max()
    return int((ev > tol).sum()), int((ev < -tol).sum()), int((np.abs(ev) <= tol).sum())


def extract_ldlt(S, F):
    """Dense (L, D, p) with A[p, p] = L D L^T (src/reference/ldlt.jl:553-588)."""
    n = S.n
    piv = F["piv"].astype(np.int64)
    pos = np.empty(n, dtype=np.int64)
    pos[piv - 1] = np.arange(n)
    L = np.zeros((n, n))
    for s in range(1, S.ns + 1):
        rows = S.snrows(s) - 1
        f = rows.shape[0]
        w = S.front_ncols[s - 1]
        c0 = S.super_ptr[s - 1] - 1
        base = S.front_ptr[s - 1] - 1
        P = F["factor"][base:base + f * w].reshape(w, f).T
        for j in range(w):
            for i in range(j, f):
                r = c0 + i if i < w else pos[rows[i]]
                L[r, c0 + j] = P[i, j]
    D = np.diag(F["d"][:n].copy())
    for k in range(n):
        if F["kind"][k] == 2:
            D[k + 1, k] = D[k, k + 1] = F["d"][n + k]
    p = S.perm[piv - 1] - 1
    return L, D, p


def extract_lu(S, F):
    """Dense (L, D, U, p, q) with A[p, q] = L D U (src/reference/lu.jl:244-275)."""
    n = S.n
    piv = F["piv"].astype(np.int64)
    ipiv = np.empty(n, dtype=np.int64)
    ipiv[piv - 1] = np.arange(n)
    L = np.zeros((n, n))
    U = np.zeros((n, n))
    for s in range(1, S.ns + 1):
        rows = S.snrows(s) - 1
        f = rows.shape[0]
        w = S.front_ncols[s - 1]
        c0 = S.super_ptr[s - 1] - 1
        base = S.front_ptr[s - 1] - 1
        PL = F["factor"][base:base + f * w].reshape(w, f).T
        PU = F["ufactor"][base:base + f * w].reshape(w, f).T
        for j in range(w):
            g = c0 + j
            for i in range(j, f):
                li = 1.0 if i == j else PL[i, j]
                ui = 1.0 if i == j else PU[i, j]
                L[c0 + i if i < w else ipiv[rows[i]], g] = li
                U[g, c0 + i if i < w else rows[i]] = ui
    p = S.perm[piv - 1] - 1
    q = S.perm - 1
    return L, F["d"][:n], U, p, q
