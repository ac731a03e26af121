"""Enumerations behind the cuDSS/CUDSS.jl strings, with the integer values of
src/types.jl:17-96 (the cuDSS C enums), which are also the values the C ABI takes."""

STRUCTURES = {"G": 0, "S": 1, "H": 2, "SPD": 3, "HPD": 4}
VIEWS = {"F": 0, "L": 1, "U": 2}
INDEX = {"Z": 0, "O": 1}

PHASES = {
    "reordering": 1,
    "symbolic_factorization": 2,
    "analysis": 3,
    "factorization": 4,
    "refactorization": 8,
    "solve_fwd_perm": 16,
    "solve_fwd": 32,
    "solve_fwd_schur": 48,
    "solve_diag": 64,
    "solve_bwd": 128,
    "solve_bwd_schur": 384,
    "solve_bwd_perm": 256,
    "solve_refinement": 512,
    "solve": 1008,
}

PIVOT_TYPES = {"A": 0, "N": 1, "C": 2, "R": 3, "D": 4, "L": 5, "B": 6}

PIVOT_KIND_1X1 = 1
PIVOT_KIND_2X2_FIRST = 2
PIVOT_KIND_2X2_SECOND = 3
PIVOT_KIND_PERTURBED = 4

# per-front statistics (src/symbolic/maps.jl:324, src/reference/ldlt.jl:25-30)
FRONT_STATS_FIELDS = 6
STAT_NPOS, STAT_NNEG, STAT_NZERO, STAT_NPERTURBED, STAT_N2X2, STAT_INFO = range(6)


def structure_code(s):
    if isinstance(s, int):
        return s
    if s not in STRUCTURES:
        raise ValueError(f'structure must be "G", "S", "H", "SPD" or "HPD", got {s!r}')
    return STRUCTURES[s]


def view_code(v):
    if isinstance(v, int):
        return v
    if v not in VIEWS:
        raise ValueError(f"view must be 'L', 'U' or 'F', got {v!r}")
    return VIEWS[v]
