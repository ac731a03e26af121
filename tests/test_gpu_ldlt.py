"""GPU LDL^T with diagonal / Bunch-Kaufman pivoting (C ABI of libsds_hip) against the CPU oracle.

The device kernels perform the CPU reference's floating-point operations in the same order
(src/reference/ldlt.jl:121-454), so in FP64 the pivot order, the pivot kinds, D, the per-front
statistics and the factor panels must be bitwise equal to the oracle's; the solve must match the
oracle's solve (test/test_numeric_ldlt.jl:48-57 asks for exact piv/kind/stats and panels within
100 eps growth; this build asks for bitwise). Inertia is also checked against numpy's
eigenvalues on the small cases.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sparsedirectsolver as sds
from generators import (eigen_inertia, kkt_interleaved_perm, kkt_matrix, random_spd, random_symindef,
                        random_symmetric)
from sparsedirectsolver import matrices as M
from sparsedirectsolver.solver import numeric_diag, numeric_factor, numeric_pivots, numeric_stats

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _solver(A, perm=None, amalgamation=(32, 0.25, 8), pivot_type="A", threshold=0.01, eps=None, sign=None,
            fp=None, structure="S"):
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), structure, "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    sds.setparam(s, "amalgamation", amalgamation)
    sds.setparam(s, "pivot_type", pivot_type)
    sds.setparam(s, "pivot_threshold", threshold)
    if eps is not None:
        sds.setparam(s, "pivot_epsilon", eps)
    if sign is not None:
        sds.setparam(s, "pivot_sign", sign)
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    return s, nz


def _rng(seed):
    return np.random.default_rng(seed)


CASES = [
    # name, matrix, perm, amalgamation, pivot_type, threshold
    ("symindef_80_amd", lambda: random_symindef(80, 0.08, _rng(1)), None, (32, 0.25, 8), "A", 0.01),
    ("symindef_300_amd_big", lambda: random_symindef(300, 0.03, _rng(2)), None, (64, 0.5, 8), "A", 0.01),
    ("symindef_200_diag", lambda: random_symindef(200, 0.04, _rng(3)), None, (32, 0.25, 8), "D", 0.1),
    ("spd_150_none", lambda: random_spd(150, 0.05, _rng(4)), None, (32, 0.25, 8), "N", 0.01),
    ("kkt_120_40_amd", lambda: kkt_matrix(120, 40, 1e-8, _rng(5)), None, (32, 0.25, 8), "B", 0.01),
    ("kkt_240_120_interleaved", lambda: kkt_matrix(240, 120, 1e-8, _rng(6)), kkt_interleaved_perm(240, 120),
     (128, 0.5, 8), "B", 0.01),
    ("kkt_400_200_delta0_amd", lambda: kkt_matrix(400, 200, 0.0, _rng(7)), None, (96, 0.4, 8), "B", 0.5),
    # the configuration's KKT generator (random H, ~10 nnz per row): wide large fronts, many
    # launch pairs per level, 2x2 pivots and the best-1x1 fallback
    ("kkt_config_1500_750_amd", lambda: M.kkt_matrix(1500, 750), None, (512, 0.25, 8), "A", 0.01),
    ("symrand_150_zero_diag", lambda: random_symmetric(150, 0.05, _rng(5), 0.0), None, (32, 0.25, 8), "A", 0.01),
    ("symrand_400_bk_big", lambda: random_symmetric(400, 0.02, _rng(5), 0.01), None, (128, 0.5, 8), "A", 0.01),
    ("symrand_250_diag_big", lambda: random_symmetric(250, 0.04, _rng(5), 0.1), None, (32, 0.25, 8), "D", 0.01),
    ("symindef_160_one_front", lambda: random_symindef(160, 0.05, _rng(8)), None, (100000, 1.0, 100000), "A", 0.01),
]


@pytest.mark.parametrize("name,make,perm,amal,ptype,thr", CASES, ids=[c[0] for c in CASES])
def test_ldlt_bitwise_oracle(name, make, perm, amal, ptype, thr):
    import torch
    A = make()
    n = A.shape[0]
    print(f"case {name} n={n}", flush=True)
    s, nz = _solver(A, perm, amal, ptype, thr)
    S = s.symbolic
    sds.execute("factorization", s)
    ref = oracle.ldlt(S, nz, pivot_type=ptype, pivot_threshold=thr, eps=1e-13)
    assert sds.getparam(s, "info") == ref["info"]
    piv, kind = numeric_pivots(s)
    assert np.array_equal(piv, ref["piv"]), "pivot order differs"
    assert np.array_equal(kind, ref["kind"]), "pivot kinds differ"
    assert np.array_equal(numeric_diag(s), ref["d"]), "D differs"
    st, tot = numeric_stats(s)
    assert np.array_equal(st, ref["stats"])
    got = numeric_factor(s)
    assert np.array_equal(got, ref["factor"]), f"panels differ by {np.abs(got - ref['factor']).max():.3e}"
    npos, nneg = sds.getparam(s, "inertia")
    if n <= 500 and ref["info"] == 0 and ptype != "N":
        ip, ineg, iz = eigen_inertia(A)
        if iz == 0:
            assert (npos, nneg) == (ip, ineg)
    # determinism
    sds.execute("refactorization", s)
    assert np.array_equal(numeric_factor(s), got)
    # solve against the oracle's solve on the same factor
    B = torch.tensor(_rng(666).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    Xr = oracle.solve_ldlt(S, ref, Bh)
    # the sweeps sum in another order than the oracle's: compare backward errors (the solutions
    # of the ill-conditioned KKT cases differ by about cond(A) eps)
    assert _berr(A, Xh, Bh) <= 10 * _berr(A, Xr, Bh) + 1e-14


def _berr(A, X, B):
    an = abs(A).sum(axis=1).max()
    return float(np.max(np.abs(B - A @ X).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))))


def test_ldlt_sub_phases_equal_solve():
    import torch
    A = random_symindef(120, 0.06, _rng(11))
    s, _ = _solver(A)
    sds.execute("factorization", s)
    B = torch.tensor(_rng(1).standard_normal((120, 3)), dtype=torch.float64, device="cuda")
    X1 = torch.zeros_like(B)
    sds.execute("solve", s, X1, B)
    X2 = torch.zeros_like(B)
    for ph in ("solve_fwd_perm", "solve_fwd", "solve_diag", "solve_bwd", "solve_bwd_perm"):
        sds.execute(ph, s, X2, B)
    assert torch.equal(X1, X2)


def test_ldlt_perturbation_and_sign():
    """The zero (1,1)-block of a KKT matrix with delta = 0 eliminated first with 'N' pivoting:
    every zero pivot is perturbed to +-eps with the requested sign; with eps = 0 the first
    zero pivot is reported through info."""
    nh, nj = 60, 30
    K = kkt_matrix(nh, nj, 0.0, _rng(12))
    n = nh + nj
    sign = np.zeros(n, dtype=np.int8)
    sign[nh:] = -1
    perm = np.concatenate([np.arange(nh + 1, n + 1), np.arange(1, nh + 1)])  # the duals first
    s, nz = _solver(K, perm, (32, 0.25, 8), "N", 0.01, eps=1e-10, sign=sign)
    sds.execute("factorization", s)
    ref = oracle.ldlt(s.symbolic, nz, pivot_type="N", pivot_threshold=0.01, eps=1e-10, pivot_sign=sign)
    piv, kind = numeric_pivots(s)
    assert np.array_equal(kind, ref["kind"]) and np.array_equal(numeric_diag(s), ref["d"])
    assert np.array_equal(numeric_stats(s)[0], ref["stats"])
    s0, nz0 = _solver(K, perm, (32, 0.25, 8), "N", 0.01, eps=0.0)
    sds.execute("factorization", s0)
    ref0 = oracle.ldlt(s0.symbolic, nz0, pivot_type="N", pivot_threshold=0.01, eps=0.0)
    assert ref0["info"] > 0
    assert sds.getparam(s0, "info") == ref0["info"]


def test_ldlt_fp32_refinement():
    """FP32 factor (pivot decisions on FP32 values) + FP64 refinement on a KKT system."""
    import torch
    K = kkt_matrix(300, 120, 1e-8, _rng(13))
    n = K.shape[0]
    s, _ = _solver(K, None, (64, 0.4, 8), "B", 0.01, fp="Float32")
    sds.execute("factorization", s)
    sds.setparam(s, "ir_n_steps", 6)
    B = torch.tensor(_rng(2).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    berr = _berr(K, X.cpu().numpy(), B.cpu().numpy())
    assert berr <= 1e-10, berr


PANEL_CASES = [c for c in CASES if c[0] in ("symindef_300_amd_big", "symindef_200_diag", "spd_150_none",
                                            "kkt_400_200_delta0_amd", "kkt_config_1500_750_amd",
                                            "symrand_400_bk_big", "symrand_250_diag_big", "symindef_160_one_front")]


@pytest.mark.parametrize("fp", [None, "Float32"])
@pytest.mark.parametrize("name,make,perm,amal,ptype,thr", PANEL_CASES, ids=[c[0] for c in PANEL_CASES])
def test_pivot_panels_bitwise_as_one_pivot_per_launch(name, make, perm, amal, ptype, thr, fp):
    """The large fronts in panels of deferred updates against one pivot per pair of launches (the
    default): pivot order, kinds, D, statistics and panels bitwise the same, in FP64 and in FP32
    (whose decisions read FP32 values): every entry receives the same operations in the same
    order, only later. Also against the CPU oracle in FP64."""
    A = make()
    print(f"case {name} n={A.shape[0]} {fp or 'Float64'}", flush=True)
    s, nz = _solver(A, perm, amal, ptype, thr, fp=fp)
    out = {}
    for pp in (0, 1):
        sds.set_tiles(s, dict(extend_add_rows=1024, cb_k_chunk=32, trailing_split=0, step_two_per_cu=0,
                              pivot_panels=pp))
        sds.execute("factorization", s)
        out[pp] = (numeric_pivots(s), numeric_diag(s), numeric_stats(s)[0], numeric_factor(s),
                   sds.getparam(s, "info"))
    (p0, k0), d0, st0, f0, i0 = out[0]
    (p1, k1), d1, st1, f1, i1 = out[1]
    assert i0 == i1
    assert np.array_equal(p0, p1) and np.array_equal(k0, k1), "pivots differ"
    assert np.array_equal(d0, d1), "D differs"
    assert np.array_equal(st0, st1)
    assert np.array_equal(f0, f1), f"panels differ by {np.abs(f0.astype(float) - f1).max():.3e}"
    if fp is None:
        ref = oracle.ldlt(s.symbolic, nz, pivot_type=ptype, pivot_threshold=thr, eps=1e-13)
        assert np.array_equal(p1, ref["piv"]) and np.array_equal(k1, ref["kind"])
        assert np.array_equal(f1, ref["factor"]) and np.array_equal(d1, ref["d"])
