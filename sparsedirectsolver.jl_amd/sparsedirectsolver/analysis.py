"""The host symbolic analysis (``Symbolic``) as numpy arrays.

Mirrors ``symbolic_analysis`` of SparseDirectSolver.jl (src/symbolic/maps.jl:255-268):
the arrays keep the reference's names and its 1-based convention, so that they
are exactly what the Julia side would pass through ``ccall``. The work is done
by ``libsds_analysis.so`` (C++).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._native import analysis_lib
from .enums import structure_code, view_code

_FIELDS = ("perm", "iperm", "parent", "counts", "super_ptr", "col2sn", "snparent", "snpost", "rowptr", "rowval",
           "front_ptr", "front_nrows", "front_ncols", "child_ptr", "child_list", "relind_ptr", "relind", "amap",
           "amap_ptr", "amap_src", "full_rowptr", "full_colval", "full_src", "etree_perm", "pairs")

# pivot_pairs settings (src/types.jl:196-203)
PIVOT_PAIRS = {"none": 0, "default": 1, "all": 2}
# matching_alg spellings (src/types.jl:143-146): MC64 jobs; "default" / "none" no matching
MATCHING_JOBS = {"algo1": 1, "algo2": 2, "algo3": 3, "algo4": 4, "algo5": 5, "algo6": 5}

# "reordering_alg" spellings (src/types.jl:100-107) and what this analysis does for them
# (src/types.jl:98-113): "default" the automatic minimum degree / nested dissection choice, "algo1" and
# "algo2" (COLAMD variants) fall back to the symmetric-pattern minimum degree as in the reference
ORDERINGS = {"default": 6, "algo1": 3, "algo2": 3, "algo3": 3, "algo4": 4, "algo5": 0, "natural": 0, "user": 1}


@dataclass
class Symbolic:
    """Result of the analysis of one pattern: 1-based int64 maps plus sizes."""
    n: int
    nnz: int
    ns: int
    structure: str
    view: str
    nnz_stored: int
    nnz_L: int
    flops: float          # stored-panel flops, sum over panel columns of rows^2 (src/symbolic/supernodes.jl:336-339)
    flops_true: float     # sum_j c_j^2 over the true column counts (src/symbolic/etree.jl:153-160)
    maps: dict = field(default_factory=dict)
    nschur: int = 0       # Schur complement mode: the size of the last supernode, the Schur block
    pair_rounds: int = 0  # rounds of the pivot_pairs = "default" search that added pairs

    @property
    def pivot_pairs(self):
        """The 2x2 pivot pairs the ordering used, as (partner, candidate) rows, 1-based."""
        return self.maps["pairs"].reshape(-1, 2)

    def __getattr__(self, name):
        maps = self.__dict__.get("maps")
        if maps is not None and name in maps:
            return maps[name]
        raise AttributeError(name)

    @property
    def factor_len(self):
        return int(self.maps["front_ptr"][-1] - 1)

    def snrows(self, s):
        """Rows (1-based, supernodal numbering) of supernode s (1-based)."""
        return self.rowval[self.rowptr[s - 1] - 1:self.rowptr[s] - 1]

    def sncols(self, s):
        return np.arange(self.super_ptr[s - 1], self.super_ptr[s])


def grid_nd_perm(nx, ny=1, nz=1, leaf=8):
    """Geometric nested dissection of an nx x ny x nz grid (natural row-major
    numbering), as a 1-based permutation: separators (the middle plane of the
    longest side) after both halves."""
    lib = analysis_lib()
    perm = np.empty(nx * ny * nz, dtype=np.int64)
    rc = lib.sdsa_grid_nd(nx, ny, nz, leaf, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc != 0:
        raise RuntimeError("grid nested dissection failed")
    return perm


def symbolic_analysis(rowptr, colval, n, structure, view="F", index="O", perm=None, ordering="default",
                      amalgamation=(32, 0.25, 8), use_superpanels=1, schur=None, values=None, pivot_pairs="none",
                      pair_tolerance=1e-6, pivot_threshold=0.01, fixed_pairs=None):
    """Run the analysis of the CSR pattern (rowptr, colval) of an n x n matrix.

    `perm` (1-based, length n) is the user permutation (``user_perm``); otherwise
    `ordering` selects "algo3" (approximate minimum degree, also the default here)
    or "algo5" (natural). `schur` (n flags 0/1, ``user_schur_indices`` of Schur complement mode,
    src/numeric/schur.jl:1-10): the flagged rows and columns go last, in increasing original
    order, and form the last supernode, which the numeric phase assembles without factoring.
    2x2 pivot pairs of "S"/"H" (src/symbolic/pairs.jl, with the AMD ordering only): `values` (the
    first member's nzval, host) and `pivot_pairs` "default" (pair the structurally zero pivots,
    partners down to `pivot_threshold` of the row maximum) or "all" (every candidate whose diagonal
    is absent or at most `pair_tolerance` times its row maximum); `fixed_pairs` ((partner,
    candidate) rows, 1-based, e.g. from `matching_pairs`) replace the search."""
    lib = analysis_lib()
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    cv = np.ascontiguousarray(colval, dtype=np.int64)
    if rp.shape[0] != n + 1:
        raise ValueError(f"length(rowptr) = {rp.shape[0]} does not match n + 1 = {n + 1}")
    base = 0 if index in ("Z", 0) else 1
    p = None
    code = ORDERINGS.get(ordering)
    if code is None:
        raise ValueError(f"reordering_alg {ordering!r} is not available in this analysis")
    if perm is not None:
        p = np.ascontiguousarray(perm, dtype=np.int64)
        if p.shape[0] != n:
            raise ValueError("user_perm must have length n")
        code = 1
    P = ctypes.POINTER(ctypes.c_int64)
    flags = None
    if schur is not None:
        flags = np.ascontiguousarray(schur, dtype=np.int8)
        if flags.shape != (n,) or np.any((flags != 0) & (flags != 1)):
            raise ValueError("user_schur_indices must hold n flags 0/1")
    vals = None
    if values is not None and pivot_pairs != "none":
        vals = np.ascontiguousarray(values, dtype=np.float64)
        if vals.shape[0] < cv.shape[0]:
            raise ValueError(f"nzval has {vals.shape[0]} entries, the pattern {cv.shape[0]}")
    if pivot_pairs not in PIVOT_PAIRS:
        raise ValueError(f"pivot_pairs must be one of {tuple(PIVOT_PAIRS)}")
    fp = None
    if fixed_pairs is not None and len(fixed_pairs):
        fp = np.ascontiguousarray(np.asarray(fixed_pairs, dtype=np.int64).reshape(-1, 2))
    h = lib.sdsa_analyze(n, rp.ctypes.data_as(P), cv.ctypes.data_as(P), structure_code(structure),
                         view_code(view), base, p.ctypes.data_as(P) if p is not None else None, code,
                         int(amalgamation[0]), float(amalgamation[1]), int(amalgamation[2]), int(use_superpanels),
                         flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)) if flags is not None else None,
                         vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if vals is not None else None,
                         PIVOT_PAIRS[pivot_pairs], float(pair_tolerance), float(pivot_threshold),
                         fp.ctypes.data_as(P) if fp is not None else None, 0 if fp is None else fp.shape[0])
    if not h:
        raise ValueError(lib.sdsa_last_error().decode())
    try:
        maps = {}
        for name in _FIELDS:
            ln = lib.sdsa_field_len(h, name.encode())
            a = np.empty(ln, dtype=np.int64)
            if ln > 0:
                lib.sdsa_field_copy(h, name.encode(), a.ctypes.data_as(P))
            maps[name] = a
        sc = [lib.sdsa_scalar(h, k) for k in range(9)]
    finally:
        lib.sdsa_free(h)
    return Symbolic(n=int(sc[0]), nnz=int(sc[1]), ns=int(sc[2]), structure=structure, view=view,
                    nnz_stored=int(sc[3]), nnz_L=int(sc[4]), flops=sc[5], flops_true=sc[6], maps=maps,
                    nschur=int(sc[7]), pair_rounds=int(sc[8]))


@dataclass
class Matching:
    """Result of `compute_matching` (src/matching/mc64.jl:25-40): row i matched to column perm[i]
    (1-based, completed to a permutation), which rows have a matched entry, the row and column
    scalings (equal for the symmetric structures), the MC64 job, and whether only the symmetric
    scaling D A D applies (the structures other than "G")."""
    job: int
    perm: np.ndarray
    matched: np.ndarray
    rscale: np.ndarray
    cscale: np.ndarray
    symmetric: bool


def compute_matching(rowptr, colval, nzval, n, structure, view="F", index="O", job=5):
    """MC64-style matching and scaling of the CSR matrix (src/matching/mc64.jl:370-392), computed by
    libsds_analysis: job 1 maximum cardinality, 2 and 3 the bottleneck, 4 the largest sum, 5 the
    largest product with the scalings (|r_i a_ij c_j| <= 1, equality on the matched entries); the
    symmetric structures get d_i = sqrt(r_i c_i) of the full matrix's matching."""
    lib = analysis_lib()
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    cv = np.ascontiguousarray(colval, dtype=np.int64)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    base = 0 if index in ("Z", 0) else 1
    perm = np.empty(n, dtype=np.int64)
    matched = np.empty(n, dtype=np.int8)
    rs = np.empty(n)
    cs = np.empty(n)
    P = ctypes.POINTER(ctypes.c_int64)
    D = ctypes.POINTER(ctypes.c_double)
    rc = lib.sdsa_matching(n, rp.ctypes.data_as(P), cv.ctypes.data_as(P), nz.ctypes.data_as(D),
                           structure_code(structure), view_code(view), base, int(job), perm.ctypes.data_as(P),
                           matched.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), rs.ctypes.data_as(D),
                           cs.ctypes.data_as(D))
    if rc != 0:
        raise ValueError(lib.sdsa_last_error().decode())
    return Matching(int(job), perm, matched.astype(bool), rs, cs, structure != "G")


def matching_pairs(m, rowptr, colval, nzval, n, structure, view="F", index="O", pivot_pairs="default",
                   pair_tolerance=1e-6):
    """The 2x2 pivot pairs (partner, candidate; 1-based rows) of "S"/"H" from the cycles of the
    symmetric matching `m` (src/matching/mc64.jl:413-461)."""
    lib = analysis_lib()
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    cv = np.ascontiguousarray(colval, dtype=np.int64)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    base = 0 if index in ("Z", 0) else 1
    out = np.empty(2 * n, dtype=np.int64)
    P = ctypes.POINTER(ctypes.c_int64)
    D = ctypes.POINTER(ctypes.c_double)
    mt = np.ascontiguousarray(m.matched, dtype=np.int8)
    d = np.ascontiguousarray(m.rscale, dtype=np.float64)
    pm = np.ascontiguousarray(m.perm, dtype=np.int64)
    k = lib.sdsa_matching_pairs(n, rp.ctypes.data_as(P), cv.ctypes.data_as(P), nz.ctypes.data_as(D),
                                structure_code(structure), view_code(view), base, pm.ctypes.data_as(P),
                                mt.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), d.ctypes.data_as(D),
                                PIVOT_PAIRS[pivot_pairs], float(pair_tolerance), out.ctypes.data_as(P))
    if k < 0:
        raise ValueError(lib.sdsa_last_error().decode())
    return out[:2 * k].reshape(-1, 2)


def entry_scaling(m, rowptr, colval, n, index="O"):
    """r_i c_j of every stored entry (i, j) of the CSR pattern, aligned with colval
    (src/matching/mc64.jl:463-470)."""
    base = 0 if index in ("Z", 0) else 1
    rp = np.asarray(rowptr, dtype=np.int64) - base
    cv = np.asarray(colval, dtype=np.int64) - base
    rows = np.repeat(np.arange(n), np.diff(rp))
    return m.rscale[rows] * m.cscale[cv]


def full_maps(rowptr, colval, n, structure, view="F", index="O"):
    """The full-pattern map of the refinement residual of a "G" matrix (the stored rows, then
    the rows of its transpose), 1-based, as the analysis builds it (src/symbolic/pattern.jl:243-290):
    used for the residual of A when the analysis ran on a matched matrix whose columns were
    relabeled."""
    if structure != "G":
        raise ValueError("full_maps rebuilds the map of a \"G\" matrix only")
    base = 0 if index in ("Z", 0) else 1
    rp = np.asarray(rowptr, dtype=np.int64) - base
    cv = np.asarray(colval, dtype=np.int64) - base
    nnz = int(rp[n])
    rows = np.repeat(np.arange(n), np.diff(rp))
    src = np.arange(nnz)
    o = np.lexsort((src, cv, rows))
    fr = np.zeros(2 * n + 1, dtype=np.int64)
    np.add.at(fr, rows[o] + 1, 1)
    fc = [cv[o] + 1]
    fs = [src[o] + 1]
    ot = o[np.argsort(cv[o], kind="stable")]
    np.add.at(fr, n + cv[ot] + 1, 1)
    fc.append(rows[ot] + 1)
    fs.append(src[ot] + 1)
    fr[0] = 1
    fr = np.cumsum(fr)
    return fr, np.concatenate(fc), np.concatenate(fs)


def merge_symbolic(parts):
    """The analysis of a block-diagonal matrix from its diagonal blocks' analyses: the members'
    arrays concatenated with their column, supernode, factor, row-list and value offsets, so
    that each member keeps exactly its own supernodes, fronts and maps (the elimination forest
    is the union of the members' trees). The members must share structure and view."""
    if not parts:
        raise ValueError("merge_symbolic needs at least one analysis")
    st, vw = parts[0].structure, parts[0].view
    if any(p.structure != st or p.view != vw for p in parts):
        raise ValueError("the members of a batch must share structure and view")
    if any(p.nschur for p in parts):
        raise ValueError("Schur complement mode is for single matrices")
    gen = st == "G"
    cat = {k: [] for k in _FIELDS}
    N = NS = NZ = FL = RV = CL = RI = AS = 0
    full_m, full_t = [], []  # (rowptr deltas, colval, src) of M rows and M^T rows
    for S in parts:
        m = S.maps
        cat["perm"].append(m["perm"] + N)
        cat["iperm"].append(m["iperm"] + N)
        cat["parent"].append(np.where(m["parent"] > 0, m["parent"] + N, 0))
        cat["counts"].append(m["counts"])
        cat["super_ptr"].append(m["super_ptr"][:-1] + N)
        cat["col2sn"].append(m["col2sn"] + NS)
        cat["snparent"].append(np.where(m["snparent"] > 0, m["snparent"] + NS, 0))
        cat["snpost"].append(m["snpost"] + NS)
        cat["rowptr"].append(m["rowptr"][:-1] + RV)
        cat["rowval"].append(m["rowval"] + N)
        cat["front_ptr"].append(m["front_ptr"][:-1] + FL)
        cat["front_nrows"].append(m["front_nrows"])
        cat["front_ncols"].append(m["front_ncols"])
        cat["child_ptr"].append(m["child_ptr"][:-1] + CL)
        cat["child_list"].append(m["child_list"] + NS)
        cat["relind_ptr"].append(m["relind_ptr"][:-1] + RI)
        cat["relind"].append(m["relind"])
        a = m["amap"]
        cat["amap"].append(np.where(a != 0, np.sign(a) * (np.abs(a) + FL), 0))
        cat["amap_ptr"].append(m["amap_ptr"][:-1] + AS)
        cat["amap_src"].append(m["amap_src"] + NZ)
        cat["etree_perm"].append(m["etree_perm"] + N)
        rp, cv, src = m["full_rowptr"], m["full_colval"], m["full_src"]
        srcs = np.sign(src) * (np.abs(src) + NZ)
        rows_m = S.n
        full_m.append((np.diff(rp[:rows_m + 1]), cv[:rp[rows_m] - 1] + N, srcs[:rp[rows_m] - 1]))
        if gen:
            full_t.append((np.diff(rp[rows_m:]), cv[rp[rows_m] - 1:] + N, srcs[rp[rows_m] - 1:]))
        N += S.n
        NS += S.ns
        NZ += S.nnz
        FL += S.factor_len
        RV += m["rowval"].shape[0]
        CL += m["child_list"].shape[0]
        RI += m["relind"].shape[0]
        AS += m["amap_src"].shape[0]
    maps = {k: np.concatenate(v) if v else np.zeros(0, dtype=np.int64) for k, v in cat.items()
            if k not in ("full_rowptr", "full_colval", "full_src")}
    for k, last in (("super_ptr", N + 1), ("rowptr", RV + 1), ("front_ptr", FL + 1), ("child_ptr", CL + 1),
                    ("relind_ptr", RI + 1), ("amap_ptr", AS + 1)):
        maps[k] = np.concatenate([maps[k], [last]]).astype(np.int64)
    blocks = full_m + full_t
    counts = np.concatenate([b[0] for b in blocks])
    maps["full_rowptr"] = np.concatenate([[1], 1 + np.cumsum(counts)]).astype(np.int64)
    maps["full_colval"] = np.concatenate([b[1] for b in blocks]).astype(np.int64)
    maps["full_src"] = np.concatenate([b[2] for b in blocks]).astype(np.int64)
    return Symbolic(n=N, nnz=NZ, ns=NS, structure=st, view=vw, nnz_stored=sum(p.nnz_stored for p in parts),
                    nnz_L=sum(p.nnz_L for p in parts), flops=float(sum(p.flops for p in parts)),
                    flops_true=float(sum(p.flops_true for p in parts)), maps=maps)
