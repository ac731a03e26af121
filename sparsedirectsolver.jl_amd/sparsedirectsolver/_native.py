"""Locations and ctypes handles of the package's native libraries.

``libsds_analysis.so`` (host C++, the symbolic analysis) and ``libsds_hip.so``
(HIP C++ for gfx950, the numeric phase behind the C ABI of ``include/sds_hip.h``)
are built in-tree by ``__graft_entry__.build()``. Nothing here falls back to a
Python or CPU implementation: a missing library raises ``ImportError``.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ANALYSIS_SO = os.path.join(PKG_ROOT, "analysis", "libsds_analysis.so")
HIP_SO = os.path.join(PKG_ROOT, "csrc", "libsds_hip.so")

_analysis = None
_hip = None


def analysis_lib():
    global _analysis
    if _analysis is None:
        if not os.path.exists(ANALYSIS_SO):
            raise ImportError(f"{ANALYSIS_SO} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(ANALYSIS_SO)
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.sdsa_analyze.restype = ctypes.c_void_p
        f64p = ctypes.POINTER(ctypes.c_double)
        i8p = ctypes.POINTER(ctypes.c_int8)
        lib.sdsa_analyze.argtypes = [ctypes.c_int64, i64p, i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p,
                                     ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, ctypes.c_int,
                                     i8p, f64p, ctypes.c_int, ctypes.c_double, ctypes.c_double, i64p, ctypes.c_int64]
        lib.sdsa_matching.restype = ctypes.c_int
        lib.sdsa_matching.argtypes = [ctypes.c_int64, i64p, i64p, f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, i64p, i8p, f64p, f64p]
        lib.sdsa_matching_pairs.restype = ctypes.c_int64
        lib.sdsa_matching_pairs.argtypes = [ctypes.c_int64, i64p, i64p, f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            i64p, i8p, f64p, ctypes.c_int, ctypes.c_double, i64p]
        lib.sdsa_last_error.restype = ctypes.c_char_p
        lib.sdsa_free.argtypes = [ctypes.c_void_p]
        lib.sdsa_field_len.restype = ctypes.c_int64
        lib.sdsa_field_len.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.sdsa_field_copy.restype = ctypes.c_int
        lib.sdsa_field_copy.argtypes = [ctypes.c_void_p, ctypes.c_char_p, i64p]
        lib.sdsa_scalar.restype = ctypes.c_double
        lib.sdsa_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.sdsa_grid_nd.restype = ctypes.c_int
        lib.sdsa_grid_nd.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, i64p]
        _analysis = lib
    return _analysis


def hip_lib():
    """The numeric library. Raises ImportError when it is not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise ImportError(f"{HIP_SO} is not built; run __graft_entry__.build()")
        from . import _abi
        _hip = _abi.bind(ctypes.CDLL(HIP_SO))
    return _hip
