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
           "amap_ptr", "amap_src", "full_rowptr", "full_colval", "full_src", "etree_perm")

# "reordering_alg" spellings (src/types.jl:100-107) and what this analysis does for them
ORDERINGS = {"default": 3, "algo3": 3, "algo5": 0, "natural": 0, "user": 1}


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
                      amalgamation=(32, 0.25, 8), use_superpanels=1):
    """Run the analysis of the CSR pattern (rowptr, colval) of an n x n matrix.

    `perm` (1-based, length n) is the user permutation (``user_perm``); otherwise
    `ordering` selects "algo3" (approximate minimum degree, also the default here)
    or "algo5" (natural)."""
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
    h = lib.sdsa_analyze(n, rp.ctypes.data_as(P), cv.ctypes.data_as(P), structure_code(structure),
                         view_code(view), base, p.ctypes.data_as(P) if p is not None else None, code,
                         int(amalgamation[0]), float(amalgamation[1]), int(amalgamation[2]), int(use_superpanels))
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
        sc = [lib.sdsa_scalar(h, k) for k in range(7)]
    finally:
        lib.sdsa_free(h)
    return Symbolic(n=int(sc[0]), nnz=int(sc[1]), ns=int(sc[2]), structure=structure, view=view,
                    nnz_stored=int(sc[3]), nnz_L=int(sc[4]), flops=sc[5], flops_true=sc[6], maps=maps)
