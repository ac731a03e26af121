"""MI355X numeric phase of SparseDirectSolver.jl's multifrontal solver.

Host API mirroring the reference (DirectSolver, execute, setparam, getparam,
update); the numeric and solve phases run in libsds_hip.so (HIP, gfx950).
"""
from .analysis import Symbolic, grid_nd_perm, symbolic_analysis
from .errors import FactorizationError, InterruptedError_, InvalidValueError, NotSupportedError
from .solver import DirectSolver, autotune, execute, getparam, set_tiles, setparam, update

__all__ = ["DirectSolver", "execute", "setparam", "getparam", "update", "autotune", "set_tiles", "Symbolic",
           "symbolic_analysis",
           "grid_nd_perm", "InvalidValueError", "NotSupportedError", "FactorizationError", "InterruptedError_"]
