"""Synthetic test and benchmark matrices as CSR arrays (scipy.sparse).

The grid generators follow test/matrices.jl:40-54 of SparseDirectSolver.jl
(Dirichlet boundary, natural row-major grid numbering, 2-D diagonal 4 and
off-diagonal -1, 3-D diagonal 6 and off-diagonal -1); the KKT generator follows
the construction of test/matrices.jl:143-158 with the sizes and distributions of
the benchmark configuration. Random streams are numpy PCG64 with fixed seeds.
"""
import numpy as np
import scipy.sparse as sp


def _lap1d(k):
    return sp.diags([-np.ones(k - 1), 2 * np.ones(k), -np.ones(k - 1)], [-1, 0, 1], format="csr")


def laplacian2d(nx, ny):
    """5-point Laplacian on an nx x ny grid (SPD, n = nx ny)."""
    return (sp.kron(sp.identity(ny), _lap1d(nx)) + sp.kron(_lap1d(ny), sp.identity(nx))).tocsr()


def laplacian3d(nx, ny, nz):
    """7-point Laplacian on an nx x ny x nz grid (SPD)."""
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    A = (sp.kron(Iz, sp.kron(Iy, _lap1d(nx))) + sp.kron(Iz, sp.kron(_lap1d(ny), Ix))
         + sp.kron(_lap1d(nz), sp.kron(Iy, Ix)))
    return A.tocsr()


def _grid_edges(dims):
    """(i, j) index pairs of the axis neighbours of a row-major grid."""
    idx = np.arange(int(np.prod(dims))).reshape(dims[::-1])
    pairs = []
    for ax in range(len(dims)):
        a = idx.take(np.arange(idx.shape[ax] - 1), axis=ax).ravel()
        b = idx.take(np.arange(1, idx.shape[ax]), axis=ax).ravel()
        pairs.append((a, b, len(dims) - 1 - ax))
    return pairs


def variable_laplacian(dims, seed=666, sigma=1.0):
    """Variable-coefficient Laplacian on a grid with cell conductivities exp(N(0, sigma^2)),
    edge coefficients the harmonic mean of their two cells, Dirichlet boundary with the
    cell's own conductivity. SPD."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    kappa = np.exp(sigma * rng.standard_normal(n))
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    deg = np.zeros(n)
    for a, b, _ in _grid_edges(dims):
        c = 2.0 * kappa[a] * kappa[b] / (kappa[a] + kappa[b])
        rows += [a, b]
        cols += [b, a]
        vals += [-c, -c]
        np.add.at(diag, a, c)
        np.add.at(diag, b, c)
        np.add.at(deg, a, 1)
        np.add.at(deg, b, 1)
    diag += (2 * len(dims) - deg) * kappa  # Dirichlet faces
    A = sp.coo_matrix((np.concatenate(vals + [diag]), (np.concatenate(rows + [np.arange(n)]),
                                                      np.concatenate(cols + [np.arange(n)]))), shape=(n, n))
    return A.tocsr()


def convection_diffusion3d(k, peclet=10.0, seed=666):
    """Nonsymmetric 3-D convection-diffusion on a k^3 grid: unit diffusion plus first-order
    upwind convection with velocity (1, 1, 1) scaled so that the cell Peclet number is `peclet`.
    M-matrix, diagonally dominant."""
    A = laplacian3d(k, k, k).tolil()
    h = 1.0 / (k + 1)
    v = peclet * 2.0 / h * h  # cell Peclet = v h / (2 * diffusion), diffusion scaled with h^2
    n = k ** 3
    idx = np.arange(n).reshape(k, k, k)
    C = sp.lil_matrix((n, n))
    for ax in range(3):
        a = idx.take(np.arange(k - 1), axis=ax).ravel()
        b = idx.take(np.arange(1, k), axis=ax).ravel()
        C[b, a] = -v  # upwind: flow towards increasing index
    C.setdiag(3 * v)
    return (A + C.tocsr()).tocsr()


def kkt_matrix(nh, nj, delta=1e-8, nnz_h_row=10, nnz_j_row=4, seed=666):
    """KKT [H J^T; J -delta I]: H random sparse symmetric with about nnz_h_row entries per
    row (off-diagonals N(0,1), diagonal = row-abs-sum + U[0,1], so H is SPD), J with
    nnz_j_row entries per row at uniform random columns, N(0,1). Inertia (nh, nj, 0)."""
    rng = np.random.default_rng(seed)
    npairs = nh * (nnz_h_row - 1) // 2
    i = rng.integers(0, nh, npairs)
    j = rng.integers(0, nh, npairs)
    keep = i != j
    i, j = i[keep], j[keep]
    v = rng.standard_normal(i.shape[0])
    H = sp.coo_matrix((np.concatenate([v, v]), (np.concatenate([i, j]), np.concatenate([j, i]))), shape=(nh, nh)).tocsr()
    H.sum_duplicates()
    d = np.asarray(abs(H).sum(axis=1)).ravel() + rng.uniform(0, 1, nh)
    H = (H + sp.diags(d)).tocsr()
    jr = np.repeat(np.arange(nj), nnz_j_row)
    jc = rng.integers(0, nh, nj * nnz_j_row)
    J = sp.coo_matrix((rng.standard_normal(jr.shape[0]), (jr, jc)), shape=(nj, nh)).tocsr()
    J.sum_duplicates()
    D = sp.diags(np.full(nj, -delta))
    K = sp.bmat([[H, J.T], [J, D]], format="csr")
    K.sort_indices()
    return K


def lower_csr(A):
    """The lower triangle (diagonal included) of a symmetric matrix, CSR with sorted columns."""
    L = sp.tril(A, format="csr")
    L.sort_indices()
    return L


def csr_arrays(A, base=1):
    """(rowptr, colval, nzval) of a scipy CSR matrix with the given index base."""
    A = A.tocsr()
    A.sort_indices()
    return (A.indptr.astype(np.int64) + base, A.indices.astype(np.int64) + base, A.data.astype(np.float64))
