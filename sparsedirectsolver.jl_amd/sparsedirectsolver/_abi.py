"""ctypes signatures of the C ABI in include/sds_hip.h."""
import ctypes

c_i32, c_i64, c_dbl, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

# the exported entry points, in the order of include/sds_hip.h
EXPORTS = ("sds_create", "sds_destroy", "sds_last_error", "sds_set_stream", "sds_upload_symbolic",
           "sds_alloc_numeric", "sds_factorize", "sds_solve", "sds_solve_noalloc", "sds_get_stats",
           "sds_get_pivots", "sds_get_diag", "sds_get_factor", "sds_memory_estimates", "sds_profile",
           "sds_profile_read", "sds_adjoint_scatter",
           "sds_autotune", "sds_graph_capture", "sds_graph_launch", "sds_graph_info", "sds_graph_destroy",
           "sds_set_ir_mode", "sds_set_tiles", "sds_set_schur", "sds_schur_export", "sds_set_matching")

# profiling kinds (SDS_K_* of include/sds_hip.h)
PROFILE_KINDS = ("front_small", "zero_scatter", "extend_add", "potrf", "trsm", "update_panel", "update_cb", "stats",
                 "permute", "fwd", "bwd", "diag", "residual", "pivot_panel", "inverse", "subtree")

SDS_SUCCESS, SDS_INVALID_VALUE, SDS_NOT_SUPPORTED, SDS_FACTORIZATION_ERROR, SDS_DEVICE_ERROR, SDS_INTERRUPTED = range(6)
SDS_R64, SDS_R32 = 0, 1


class SymbolicDesc(ctypes.Structure):
    _fields_ = [("n", c_i64), ("nnz", c_i64), ("ns", c_i64), ("structure", c_i32), ("index_bits", c_i32)] + [
        (k, vp) for k in ("perm", "super_ptr", "snparent", "rowptr", "rowval", "front_ptr", "front_nrows",
                          "front_ncols", "child_ptr", "child_list", "relind_ptr", "relind", "amap", "amap_ptr",
                          "amap_src")] + [("full_nrows", c_i64), ("full_nnz", c_i64), ("full_rowptr", vp),
                                          ("full_colval", vp), ("full_src", vp)]


class FactorParams(ctypes.Structure):
    _fields_ = [("pivot_type", c_i32), ("pivot_epsilon_alg", c_i32), ("pivot_threshold", c_dbl),
                ("pivot_epsilon", c_dbl), ("pivot_sign", vp), ("interrupt", vp)]


def bind(lib):
    P = ctypes.POINTER
    lib.sds_create.argtypes = [P(vp), vp]
    lib.sds_destroy.argtypes = [vp]
    lib.sds_last_error.argtypes = [vp]
    lib.sds_last_error.restype = ctypes.c_char_p
    lib.sds_set_stream.argtypes = [vp, vp]
    lib.sds_upload_symbolic.argtypes = [vp, P(SymbolicDesc)]
    lib.sds_alloc_numeric.argtypes = [vp, c_i32, c_i32]
    lib.sds_factorize.argtypes = [vp, vp, P(FactorParams), P(c_i64)]
    lib.sds_solve.argtypes = [vp, c_i32, c_i32, vp, c_i64, vp, c_i64, c_i32, vp, c_i32, c_dbl, P(c_i32)]
    lib.sds_solve_noalloc.argtypes = lib.sds_solve.argtypes
    lib.sds_get_stats.argtypes = [vp, vp, vp]
    lib.sds_set_schur.argtypes = [vp, c_i64]
    lib.sds_set_matching.argtypes = [vp, vp, vp, vp, vp]
    lib.sds_schur_export.argtypes = [vp, vp, c_i64]
    lib.sds_get_pivots.argtypes = [vp, vp, vp]
    lib.sds_autotune.argtypes = [vp, vp, P(FactorParams), c_i32, vp]
    lib.sds_set_tiles.argtypes = [vp, vp]
    lib.sds_graph_capture.argtypes = [vp, vp, P(FactorParams), c_i32, c_i32, vp, c_i64, vp, c_i64, c_i32, c_i32]
    lib.sds_graph_launch.argtypes = [vp]
    lib.sds_set_ir_mode.argtypes = [vp, c_i32]
    lib.sds_graph_info.argtypes = [vp, P(c_i64)]
    lib.sds_graph_destroy.argtypes = [vp]
    lib.sds_adjoint_scatter.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, c_i32, vp]
    lib.sds_get_diag.argtypes = [vp, vp]
    lib.sds_get_factor.argtypes = [vp, c_i32, vp]
    lib.sds_memory_estimates.argtypes = [vp, vp]
    lib.sds_profile.argtypes = [vp, c_i32]
    lib.sds_profile_read.argtypes = [vp, c_i32, P(c_dbl), P(c_i64), P(c_dbl), P(c_dbl)]
    for name in EXPORTS:
        if name != "sds_last_error":
            getattr(lib, name).restype = ctypes.c_int
    return lib
