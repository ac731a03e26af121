"""GPU Cholesky (C ABI of libsds_hip) against the CPU oracle.

Contract of the reference's device-vs-oracle tests (test/test_numeric_cholesky_a.jl:69-94,
test/utils.jl:230, :377): panel error max|L_gpu - L_ref| / max|L_ref| <= 100 eps, per-front
statistics equal, solution relative residual <= sqrt(eps), and a second factorization
bitwise identical to the first.
"""
import numpy as np
import pytest

import oracle
import sparsedirectsolver as sds
from sparsedirectsolver import matrices as M
from sparsedirectsolver.solver import numeric_factor, numeric_stats, profile, profile_read

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _solver(A, perm=None, amalgamation=(32, 0.25, 8), fp=None, ordering=None):
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "SPD", "L")
    if perm is not None:
        sds.setparam(s, "user_perm", perm)
    if ordering is not None:
        sds.setparam(s, "reordering_alg", ordering)
    sds.setparam(s, "amalgamation", amalgamation)
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    return s, nz


def _block_arrow(nb=2, bs=540, tail=530, link=100, seed=5):
    """nb dense bs x bs blocks whose first columns couple to `link` random columns of a dense
    tail block, diagonally dominant. In natural order every block is one bs-column front with
    `link` rows below and the tail is the root."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    n = nb * bs + tail
    rows, cols, vals = [], [], []

    def dense(o, k):
        r, c = np.tril_indices(k, -1)
        rows.extend(r + o)
        cols.extend(c + o)
        vals.extend(rng.uniform(-1, 1, r.size))

    for b in range(nb):
        dense(b * bs, bs)
        t = nb * bs + rng.choice(tail, link, replace=False)
        rows.extend(t)
        cols.extend([b * bs] * link)
        vals.extend(rng.uniform(-1, 1, link))
    dense(nb * bs, tail)
    L = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A = L + L.T
    return (A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)).tocsr()


CASES = [
    ("lap2d_10x10_natural", lambda: (M.laplacian2d(10, 10), None, (32, 0.25, 8))),
    ("lap2d_30x20_nd", lambda: (M.laplacian2d(30, 20), sds.grid_nd_perm(30, 20, 1), (32, 0.25, 8))),
    ("lap3d_12_nd", lambda: (M.laplacian3d(12, 12, 12), sds.grid_nd_perm(12, 12, 12), (32, 0.25, 8))),
    ("lap3d_16_nd_wide", lambda: (M.laplacian3d(16, 16, 16), sds.grid_nd_perm(16, 16, 16), (512, 0.25, 8))),
    ("varlap3d_14_nd", lambda: (M.variable_laplacian((14, 13, 12), seed=3), sds.grid_nd_perm(14, 13, 12),
                                (128, 0.25, 8))),
    ("lap3d_20_natural", lambda: (M.laplacian3d(20, 20, 20), None, (64, 0.25, 8))),
    # a 784-column root front: wide solve steps (more than 512 columns) and a 13-step chain,
    # element by element
    ("lap3d_28_nd_wide", lambda: (M.laplacian3d(28, 28, 28), sds.grid_nd_perm(28, 28, 28), (512, 0.25, 8))),
    # the reference's CPU-runnable configuration: 2-D 100 x 100, the build's AMD ordering
    ("lap2d_100x100_amd", lambda: (M.laplacian2d(100, 100), None, (32, 0.25, 8))),
    # fronts of 540 columns with 100 rows below and a 526-column root: a wide solve step of 512
    # columns followed by a narrow one (28 and 14 columns) that starts at column 512
    ("block_arrow_540_narrow_last_step", lambda: (_block_arrow(), np.arange(1, 1611), (512, 0.25, 8))),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_factor_matches_oracle(name, make):
    import torch
    A, perm, amal = make()
    print(f"case {name} n={A.shape[0]}", flush=True)
    s, nz = _solver(A, perm, amal)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    info, ref, st_ref = oracle.cholesky(s.symbolic, nz)
    assert info == 0
    got = numeric_factor(s)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 100 * EPS, f"panel error {err:.3e}"
    st, tot = numeric_stats(s)
    assert np.array_equal(st, st_ref)
    # determinism: a refactorization gives bitwise identical panels
    sds.execute("refactorization", s)
    assert np.array_equal(numeric_factor(s), got)
    # solve, several right-hand sides
    n = A.shape[0]
    rng = np.random.default_rng(666)
    B = torch.tensor(rng.standard_normal((n, 3)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    Xr = oracle.solve_cholesky(s.symbolic, ref, Bh)
    assert np.abs(Xh - Xr).max() / np.abs(Xr).max() <= 1e-10
    relres = np.linalg.norm(A @ Xh - Bh, axis=0) / np.maximum(np.linalg.norm(Bh, axis=0), 1)
    assert relres.max() <= np.sqrt(EPS)


def test_failed_pivot_info():
    """[4 0 0 0; 0 4 0 0; 0 0 1 2; 0 0 2 1] has its first non-positive pivot at column 4
    (test/test_reference_cholesky.jl:121-124)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(np.array([[4, 0, 0, 0], [0, 4, 0, 0], [0, 0, 1, 2], [0, 0, 2, 1]], dtype=float))
    s, nz = _solver(A, perm=np.arange(1, 5))
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 4
    info, _, _ = oracle.cholesky(s.symbolic, nz)
    assert info == 4


def test_sub_phases_equal_solve():
    import torch
    A = M.laplacian3d(10, 9, 8)
    s, nz = _solver(A, sds.grid_nd_perm(10, 9, 8), (64, 0.25, 8))
    sds.execute("factorization", s)
    n = A.shape[0]
    B = torch.tensor(np.random.default_rng(1).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    X1 = torch.zeros_like(B)
    sds.execute("solve", s, X1, B)
    X2 = torch.zeros_like(B)
    for ph in ("solve_fwd_perm", "solve_fwd", "solve_diag", "solve_bwd", "solve_bwd_perm"):
        sds.execute(ph, s, X2, B)
    assert torch.equal(X1, X2)


def test_refinement_fp64_and_fp32():
    import torch
    A = M.laplacian3d(16, 16, 16)
    perm = sds.grid_nd_perm(16, 16, 16)
    n = A.shape[0]
    B = torch.tensor(np.random.default_rng(2).standard_normal((n, 8)), dtype=torch.float64, device="cuda")
    Bh = B.cpu().numpy()
    for fp, steps, tol in ((None, 2, 1e-14), ("Float32", 6, 1e-10)):
        s, nz = _solver(A, perm, (512, 0.25, 8), fp=fp)
        sds.execute("factorization", s)
        sds.setparam(s, "ir_n_steps", steps)
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        Xh = X.cpu().numpy()
        berr = np.abs(A @ Xh - Bh).max(axis=0) / (abs(A).sum(axis=1).max() * np.abs(Xh).max(axis=0) +
                                                  np.abs(Bh).max(axis=0))
        assert berr.max() <= tol, (fp, berr.max())


@pytest.mark.parametrize("ordering", ["grid_nd", "algo4"])
def test_laplacian3d_64_full_size_backward_error(ordering):
    """The benchmark's configuration at full size (3-D 64^3, nested dissection of the grid or of
    the graph, 8 right-hand sides): size-independent checks, as the oracle would take minutes
    here. FP64 backward error <= 1e-14, inertia-free stats (every pivot positive),
    bitwise-reproducible refactorization and solve. The graph dissection gives wide fronts with
    rows below whose last solve step is narrow."""
    import torch
    k = 64
    A = M.laplacian3d(k, k, k)
    if ordering == "grid_nd":
        s, nz = _solver(A, sds.grid_nd_perm(k, k, k), (512, 0.25, 8))
    else:
        s, nz = _solver(A, None, (512, 0.25, 8), ordering=ordering)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    n = A.shape[0]
    B = torch.tensor(np.random.default_rng(666).standard_normal((n, 8)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    sds.execute("solve", s, X, B)
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    an = abs(A).sum(axis=1).max()
    berr = np.max(np.abs(Bh - A @ Xh).max(axis=0) / (an * np.abs(Xh).max(axis=0) + np.abs(Bh).max(axis=0)))
    assert berr <= 1e-14, f"backward error {berr:.2e}"
    sds.execute("refactorization", s)
    X2 = torch.zeros_like(B)
    sds.execute("solve", s, X2, B)
    assert torch.equal(X, X2)


def test_every_tile_candidate_gives_the_same_factor():
    """Every tile shape the autotuner can pick (sds_autotune's candidates, forced one at a time
    through sds_set_tiles) gives bitwise the factor of the default shape. The 56 x 56 x 40
    Laplacian in nested dissection with (512, 0.25, 8) has a step with 1,282 trailing tiles over
    its level and step launches of 1,564 workgroups, so every trailing-split threshold (256, 512,
    1024) and every two-per-CU threshold (1 to 512) switches some launch over. Its small fronts
    have from 8 to 90 rows, so each small-front kernel choice (one wave per front of at most 32
    rows with blocked or column-by-column panels above, or column by column for all) factors some;
    a solve with the same right-hand sides is bitwise the same too (the diagonal-block inverses
    formed in the front kernels or by the inverse kernel)."""
    import torch
    A = M.laplacian3d(56, 56, 40)
    s, nz = _solver(A, sds.grid_nd_perm(56, 56, 40), (512, 0.25, 8))
    base = dict(extend_add_rows=1024, cb_k_chunk=32, trailing_split=0, step_two_per_cu=0, small_front_kernel=0)
    sds.set_tiles(s, base)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    ref = numeric_factor(s)
    B = torch.tensor(np.random.default_rng(4).standard_normal((A.shape[0], 2)), dtype=torch.float64, device="cuda")
    Xref = torch.zeros_like(B)
    sds.execute("solve", s, Xref, B)
    cands = [dict(base, extend_add_rows=ea, cb_k_chunk=kc) for ea in (512, 1024, 2048, 4096) for kc in (16, 32, 64)]
    cands += [dict(base, trailing_split=rs) for rs in (256, 512, 1024)]
    cands += [dict(base, step_two_per_cu=oc) for oc in (1, 64, 128, 256, 512)]
    cands += [dict(base, trailing_split=256, step_two_per_cu=1)]
    cands += [dict(base, small_front_kernel=k) for k in (1, 2)]
    cands += [dict(base, subtree_kib=b) for b in (16, 32, 64)]
    for c in cands:
        print("tiles", c, flush=True)
        sds.set_tiles(s, c)
        sds.execute("factorization", s)
        assert sds.getparam(s, "info") == 0
        assert np.array_equal(numeric_factor(s), ref), f"tile shape {c} changes the factor"
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        assert torch.equal(X, Xref), f"tile shape {c} changes the solve"


SUBTREE_CASES = [
    # the reference's CPU-runnable configuration: 2-D 100 x 100, the build's AMD ordering
    ("lap2d_100x100_amd", lambda: (M.laplacian2d(100, 100), None, (32, 0.25, 8))),
    ("varlap3d_14_nd", lambda: (M.variable_laplacian((14, 13, 12), seed=3), sds.grid_nd_perm(14, 13, 12),
                                (128, 0.25, 8))),
    ("lap3d_20_natural", lambda: (M.laplacian3d(20, 20, 20), None, (64, 0.25, 8))),
]


@pytest.mark.parametrize("fp", [None, "Float32"])
@pytest.mark.parametrize("name,make", SUBTREE_CASES, ids=[c[0] for c in SUBTREE_CASES])
def test_subtree_kernel_matches_oracle(name, make, fp):
    """Regime A (leaf subtrees, one workgroup each, the stack in LDS) against the CPU oracle: the
    factor within 100 eps (FP64; FP32 within 1e-5 relative of the oracle's FP64 factor), the
    statistics equal, the factor and the solve bitwise those of the level kernels, and the
    subtree launches actually taking the subtrees (fewer level fronts)."""
    import torch
    A, perm, amal = make()
    s, nz = _solver(A, perm, amal, fp=fp)
    tiles = dict(extend_add_rows=1024, cb_k_chunk=32, trailing_split=0, step_two_per_cu=0, small_front_kernel=0)
    n = A.shape[0]
    B = torch.tensor(np.random.default_rng(8).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    out = {}
    for kib in (0, 16, 64):
        sds.set_tiles(s, dict(tiles, subtree_kib=kib))
        profile(s, True)
        sds.execute("factorization", s)
        prof = profile_read(s)
        profile(s, False)
        assert sds.getparam(s, "info") == 0
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        out[kib] = (numeric_factor(s), numeric_stats(s)[0], X.cpu().numpy(), prof.get("subtree", (0, 0))[1])
        print(f"{name} fp={fp} subtree {kib} KiB: {out[kib][3]} subtree launches", flush=True)
    assert out[0][3] == 0 and out[64][3] >= 1
    info, ref, st_ref = oracle.cholesky(s.symbolic, nz)
    for kib in (16, 64):
        got, st, X, _ = out[kib]
        assert np.array_equal(got, out[0][0]) and np.array_equal(st, st_ref)
        assert np.array_equal(X, out[0][2])
        err = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
        assert err <= (100 * EPS if fp is None else 1e-5), err


def test_autotune_keeps_the_factor():
    """The tile-shape autotuner picks among shapes that apply the same operations in the same
    order to every entry, except the panel block, which regroups a panel entry's updates: the
    factor after tuning is bitwise the one before when the panel block stayed 0, else within
    100 eps of it, and bitwise the factor of the same shape set by hand."""
    A = M.laplacian3d(24, 23, 22)
    s, nz = _solver(A, sds.grid_nd_perm(24, 23, 22), (512, 0.25, 8))
    sds.execute("factorization", s)
    before = numeric_factor(s)
    tiles = sds.autotune(s, reps=1)
    print("autotuned", tiles, flush=True)
    assert tiles["extend_add_rows"] in (512, 1024, 2048, 4096) and tiles["cb_k_chunk"] in (16, 32, 64)
    assert tiles["trailing_split"] in (0, 256, 512, 1024)
    assert tiles["step_two_per_cu"] in (0, 1, 64, 128, 256, 512)
    assert tiles["panel_block"] in (0, 128, 256, 512)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    after = numeric_factor(s)
    if tiles["panel_block"] == 0:
        assert np.array_equal(after, before)
    else:
        assert np.abs(after - before).max() <= 100 * EPS * np.abs(before).max()
    sds.set_tiles(s, tiles)
    sds.execute("factorization", s)
    assert np.array_equal(numeric_factor(s), after)


_LAP28 = {}


def _lap28_oracle():
    """lap3d_28 in nested dissection with (512, 0.25, 8) (a 784-column root front) and the CPU
    oracle's factor of it, computed once for the panel-block cases."""
    if not _LAP28:
        A = M.laplacian3d(28, 28, 28)
        s, nz = _solver(A, sds.grid_nd_perm(28, 28, 28), (512, 0.25, 8))
        info, ref, st_ref = oracle.cholesky(s.symbolic, nz)
        assert info == 0
        _LAP28.update(A=A, ref=ref, st=st_ref)
    return _LAP28


PANEL_BLOCK_CASES = [
    # (panel block, trailing split, two-per-CU steps, factor precision)
    (64, 0, 0, None),     # every step a block: no chain, every strip from k_big_block
    (128, 0, 0, None),    # 6 full blocks of the 784-column root: update tiles ride in the steps
    (128, 256, 1, None),  # ... or get launches of their own; the two-per-CU step build
    (256, 0, 0, None),
    (512, 0, 0, None),    # most fronts one block, the root two
    (256, 0, 0, "Float32"),
]


@pytest.mark.parametrize("pb,rs,oc,fp", PANEL_BLOCK_CASES,
                         ids=[f"pb{c[0]}_rs{c[1]}_oc{c[2]}_{c[3] or 'Float64'}" for c in PANEL_BLOCK_CASES])
def test_panel_blocks_match_oracle(pb, rs, oc, fp):
    """Large fronts factored in panel blocks (the step chain inside each block, the block's update
    of the columns to its right with K = the block width) against the CPU oracle on a front of
    784 columns: the factor within 100 eps (FP32: 1e-5 relative to the oracle's FP64 factor), the
    statistics equal, a refactorization bitwise the same, and the solve's backward error."""
    import torch
    case = _lap28_oracle()
    A = case["A"]
    s, nz = _solver(A, sds.grid_nd_perm(28, 28, 28), (512, 0.25, 8), fp=fp)
    sds.set_tiles(s, dict(extend_add_rows=1024, cb_k_chunk=32, trailing_split=rs, step_two_per_cu=oc,
                          small_front_kernel=0, subtree_kib=0, panel_block=pb))
    print(f"lap3d_28 panel block {pb} split {rs} two-per-CU {oc} {fp or 'Float64'}", flush=True)
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    got = numeric_factor(s)
    ref = case["ref"]
    err = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
    assert err <= (100 * EPS if fp is None else 1e-5), f"panel block {pb}: factor error {err:.3e}"
    assert np.array_equal(numeric_stats(s)[0], case["st"])
    sds.execute("refactorization", s)
    assert np.array_equal(numeric_factor(s), got)
    n = A.shape[0]
    B = torch.tensor(np.random.default_rng(12).standard_normal((n, 2)), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    if fp:
        sds.setparam(s, "ir_n_steps", 3)
    sds.execute("solve", s, X, B)
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    an = abs(A).sum(axis=1).max()
    berr = np.max(np.abs(Bh - A @ Xh).max(axis=0) / (an * np.abs(Xh).max(axis=0) + np.abs(Bh).max(axis=0)))
    assert berr <= (1e-14 if fp is None else 1e-10), f"backward error {berr:.2e}"


@pytest.mark.parametrize("sigma", [2.0, 4.0])
def test_ill_conditioned_solve_matches_the_oracle(sigma):
    """The sweeps multiply by inverted diagonal blocks (64 x 64, and whole 512-column steps for the
    wide fronts) where the oracle substitutes (src/reference/cholesky.jl:127-170). On a
    variable-coefficient Laplacian with conductivities exp(N(0, sigma^2)) (sigma = 4: a condition
    number of order 1e10 and more) over a 784-column root, the GPU solve's backward error must stay
    within 10x the oracle's (plus 1e-15), with and without refinement."""
    import torch
    k = 28
    A = M.variable_laplacian((k, k, k), seed=17, sigma=sigma)
    s, nz = _solver(A, sds.grid_nd_perm(k, k, k), (512, 0.25, 8))
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    info, ref, _ = oracle.cholesky(s.symbolic, nz)
    assert info == 0
    n = A.shape[0]
    Bh = np.random.default_rng(3).standard_normal((n, 2))
    B = torch.tensor(Bh, dtype=torch.float64, device="cuda")
    Xr = oracle.solve_cholesky(s.symbolic, ref, Bh)
    an = abs(A).sum(axis=1).max()

    def berr(X):
        return float(np.max(np.abs(Bh - A @ X).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(Bh).max(axis=0))))

    for steps in (0, 2):
        sds.setparam(s, "ir_n_steps", steps)
        X = torch.zeros_like(B)
        sds.execute("solve", s, X, B)
        got, want = berr(X.cpu().numpy()), berr(Xr)
        print(f"sigma {sigma} ir {steps}: GPU backward error {got:.2e}, oracle {want:.2e}", flush=True)
        assert got <= 10 * want + 1e-15, (got, want)


def test_interrupt_flag_stops_the_factorization():
    """user_host_interrupt (a one-byte flag) set before the phase: the library's level loop sees
    it before the first launch group and the phase raises InterruptedError
    (src/numeric/factorize.jl:365-379, src/logging.jl:59-63); cleared, the same solver factors."""
    import ctypes
    A = M.laplacian3d(10, 10, 10)
    s, nz = _solver(A, sds.grid_nd_perm(10, 10, 10))
    flag = ctypes.c_bool(True)
    sds.setparam(s, "user_host_interrupt", flag)
    with pytest.raises(sds.InterruptedError_):
        sds.execute("factorization", s)
    flag.value = False
    sds.execute("factorization", s)
    assert sds.getparam(s, "info") == 0
    with pytest.raises(sds.InvalidValueError):
        sds.setparam(s, "user_host_interrupt", ctypes.c_int32(0))
        sds.execute("factorization", s)
