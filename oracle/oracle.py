"""ctypes bindings of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as
the checker; the product path never imports this module.

Each function restates one function of SparseDirectSolver.jl's CPU reference
(see oracle.c for the file:line of each):
  cholesky       ref_factorize!   src/reference/cholesky.jl:44-111
  solve_cholesky ref_solve!       src/reference/cholesky.jl:127-170
  ldlt           ref_ldlt!        src/reference/ldlt.jl:121-202, _ldlt_front! :384-454
  solve_ldlt     ref_solve_ldlt!  src/reference/ldlt.jl:467-539
  lu             ref_lu!          src/reference/lu.jl:84-155, _lu_front! :163-214
  solve_lu       ref_solve_lu!    src/reference/lu.jl:283-294
  ldlt_front     _ldlt_front!     src/reference/ldlt.jl:384-454 (one dense front on its own)
  lu_front       _lu_front!       src/reference/lu.jl:163-214 (one dense front on its own)

SDS_ORACLE_LIB, when set, names another build of oracle.c to load (the mutation tests
build deliberately broken copies in a temporary folder and check that the pins fail).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("SDS_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_I64P = ctypes.POINTER(ctypes.c_int64)


class _Sym(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("ns", ctypes.c_int64),
                ("factor_len", ctypes.c_int64)] + [(k, _I64P) for k in (
                    "perm", "super_ptr", "snparent", "snpost", "rowptr", "rowval", "front_ptr", "front_nrows",
                    "front_ncols", "child_ptr", "child_list", "relind_ptr", "relind", "amap", "amap_ptr",
                    "amap_src")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.oracle_cholesky.restype = ctypes.c_int64
        L.oracle_cholesky.argtypes = [vp, vp, vp, vp]
        L.oracle_cholesky_prefix.restype = ctypes.c_int64
        L.oracle_cholesky_prefix.argtypes = [vp, vp, vp, vp, ctypes.c_int64]
        L.oracle_solve_cholesky.argtypes = [vp, vp, vp, vp, ctypes.c_int64]
        L.oracle_ldlt.restype = ctypes.c_int64
        L.oracle_ldlt.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp, vp, vp]
        L.oracle_solve_ldlt.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64]
        L.oracle_lu.restype = ctypes.c_int64
        L.oracle_lu.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp, vp, vp]
        L.oracle_solve_lu.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int]
        i64, dbl = ctypes.c_int64, ctypes.c_double
        L.oracle_ldlt_front.argtypes = [vp, i64, i64, ctypes.c_int, dbl, dbl, vp, vp, vp, vp, vp, vp]
        L.oracle_lu_front.argtypes = [vp, i64, i64, ctypes.c_int, dbl, dbl, vp, vp, vp, vp]
        L.oracle_front_begin.restype = vp
        L.oracle_front_begin.argtypes = [vp]
        L.oracle_front_assemble.argtypes = [vp, i64, vp, vp]
        L.oracle_front_keep.argtypes = [vp, i64, vp, vp]
        L.oracle_front_end.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class OracleSym:
    """The analysis arrays packed for the oracle (keeps them alive)."""

    def __init__(self, S):
        self.S = S
        self.keep = {}
        st = _Sym()
        st.n, st.nnz, st.ns, st.factor_len = S.n, S.nnz, S.ns, S.factor_len
        for name, _ in _Sym._fields_[4:]:
            a = np.ascontiguousarray(S.maps[name], dtype=np.int64)
            if a.size == 0:
                a = np.zeros(1, dtype=np.int64)
            self.keep[name] = a
            setattr(st, name, a.ctypes.data_as(_I64P))
        self.st = st

    @property
    def ptr(self):
        return ctypes.byref(self.st)


def _rhs(B, n):
    B = np.asarray(B, dtype=np.float64)
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    return np.asfortranarray(B.reshape(n, nrhs)), nrhs


def cholesky(S, nzval):
    """Returns (info, factor, stats)."""
    o = OracleSym(S)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    factor = np.zeros(S.factor_len)
    stats = np.zeros(6 * S.ns, dtype=np.int64)
    info = lib().oracle_cholesky(o.ptr, _p(nz), _p(factor), _p(stats))
    return int(info), factor, stats


def cholesky_prefix(S, nzval, q1):
    """The oracle Cholesky on the first q1 supernodes of the postorder only (a union of whole
    subtrees); used to time a bounded sample of a large workload."""
    o = OracleSym(S)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    factor = np.zeros(S.factor_len)
    stats = np.zeros(6 * S.ns, dtype=np.int64)
    info = lib().oracle_cholesky_prefix(o.ptr, _p(nz), _p(factor), _p(stats), int(q1))
    return int(info), factor, stats


def solve_cholesky(S, factor, B):
    o = OracleSym(S)
    Bf, nrhs = _rhs(B, S.n)
    X = np.zeros_like(Bf)
    lib().oracle_solve_cholesky(o.ptr, _p(np.ascontiguousarray(factor)), _p(Bf), _p(X), nrhs)
    return X.reshape(np.shape(B))


PIVOTS = {"N": 0, "D": 1, "B": 2, "A": 2, "L": 2}


def ldlt(S, nzval, pivot_type="A", pivot_threshold=0.01, eps=1e-13, pivot_sign=None):
    """Returns dict(info, factor, d, piv, kind, stats). `eps` is the effective perturbation."""
    o = OracleSym(S)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    n = S.n
    out = dict(factor=np.zeros(S.factor_len), d=np.zeros(2 * n), piv=np.zeros(n, dtype=np.int32),
               kind=np.zeros(n, dtype=np.int8), stats=np.zeros(6 * S.ns, dtype=np.int64))
    ps = None if pivot_sign is None else np.ascontiguousarray(pivot_sign, dtype=np.int8)
    out["info"] = int(lib().oracle_ldlt(o.ptr, _p(nz), PIVOTS[pivot_type], pivot_threshold, eps, _p(ps),
                                        _p(out["factor"]), _p(out["d"]), _p(out["piv"]), _p(out["kind"]),
                                        _p(out["stats"])))
    return out


def solve_ldlt(S, F, B):
    o = OracleSym(S)
    Bf, nrhs = _rhs(B, S.n)
    X = np.zeros_like(Bf)
    lib().oracle_solve_ldlt(o.ptr, _p(F["factor"]), _p(F["d"]), _p(F["piv"]), _p(F["kind"]), _p(Bf), _p(X), nrhs)
    return X.reshape(np.shape(B))


def lu(S, nzval, pivoting=True, pivot_threshold=0.01, eps=1e-13):
    o = OracleSym(S)
    nz = np.ascontiguousarray(nzval, dtype=np.float64)
    n = S.n
    out = dict(factor=np.zeros(S.factor_len), ufactor=np.zeros(S.factor_len), d=np.zeros(2 * n),
               piv=np.zeros(n, dtype=np.int32), kind=np.zeros(n, dtype=np.int8),
               stats=np.zeros(6 * S.ns, dtype=np.int64))
    out["info"] = int(lib().oracle_lu(o.ptr, _p(nz), int(pivoting), pivot_threshold, eps, _p(out["factor"]),
                                      _p(out["ufactor"]), _p(out["d"]), _p(out["piv"]), _p(out["kind"]),
                                      _p(out["stats"])))
    return out


def solve_lu(S, F, B, transpose=False):
    o = OracleSym(S)
    Bf, nrhs = _rhs(B, S.n)
    X = np.zeros_like(Bf)
    lib().oracle_solve_lu(o.ptr, _p(F["factor"]), _p(F["ufactor"]), _p(F["d"]), _p(F["piv"]), _p(Bf), _p(X), nrhs,
                          int(transpose))
    return X.reshape(np.shape(B))


def extract_L(S, factor):
    """Dense lower-triangular L (supernodal numbering) from the panels, every stored entry."""
    n = S.n
    L = np.zeros((n, n))
    for s in range(1, S.ns + 1):
        rows = S.snrows(s) - 1
        f = rows.shape[0]
        c0 = S.super_ptr[s - 1] - 1
        w = S.front_ncols[s - 1]
        base = S.front_ptr[s - 1] - 1
        P = factor[base:base + f * w].reshape(w, f).T
        for j in range(w):
            L[rows[j:], c0 + j] = P[j:, j]
    return L


def ldlt_front(F, w, pivot_type="A", pivot_threshold=0.01, eps=1e-13, sign_req=None):
    """The dense in-front LDL^T of one assembled f x f front (both triangles of the leading w x w
    block filled). Returns dict(F (factored in place, a copy), lp (0-based local pivot order),
    kind, d, e, stats). The per-column outputs are allocated f long and cut to w."""
    F = np.array(F, dtype=np.float64, order="F", copy=True)
    f = F.shape[0]
    lp = np.zeros(f, dtype=np.int64)
    kind = np.zeros(f, dtype=np.int8)
    d = np.zeros(f)
    e = np.zeros(f)
    st = np.zeros(6, dtype=np.int64)
    sr = None if sign_req is None else np.ascontiguousarray(sign_req, dtype=np.int8)
    lib().oracle_ldlt_front(_p(F), f, int(w), PIVOTS[pivot_type], float(pivot_threshold), float(eps), _p(sr),
                            _p(lp), _p(kind), _p(d), _p(e), _p(st))
    return dict(F=F, lp=lp[:w], kind=kind[:w], d=d[:w], e=e[:w], stats=st)


def lu_front(F, w, pivoting=True, pivot_threshold=0.01, eps=1e-13):
    """The dense in-front L D U of one assembled f x f front. Returns dict(F (a factored copy),
    lp (0-based local row order), kind, d, stats)."""
    F = np.array(F, dtype=np.float64, order="F", copy=True)
    f = F.shape[0]
    lp = np.zeros(f, dtype=np.int64)
    kind = np.zeros(f, dtype=np.int8)
    d = np.zeros(f)
    st = np.zeros(6, dtype=np.int64)
    lib().oracle_lu_front(_p(F), f, int(w), int(pivoting), float(pivot_threshold), float(eps), _p(lp), _p(kind),
                          _p(d), _p(st))
    return dict(F=F, lp=lp[:w], kind=kind[:w], d=d[:w], stats=st)
