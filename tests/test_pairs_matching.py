"""2x2 pivot pairs (src/symbolic/pairs.jl) and MC64-style matching and scaling
(src/matching/mc64.jl) of the host analysis, on the CPU.

The reference holds no vectors for these routines; they are pinned by what the mathematics fixes:
the optimal objective of each MC64 job against scipy's assignment and bipartite-matching solvers
(the matching itself may differ on ties), the dual feasibility of the job-5 scalings
(|r_i a_ij c_j| <= 1, equality on the matched entries), the structural zero pivots against a
brute-force structural-rank recurrence over the elimination prefixes, and the pair pattern's
invariants (partner right before its candidate, both in one supernode, the union structure).
Mutations of the restated rules are caught by these checks (see test_mutations_are_caught).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linear_sum_assignment
from scipy.sparse.csgraph import maximum_bipartite_matching

from generators import kkt_matrix, random_symmetric, random_unsym
from sparsedirectsolver import matrices as M
from sparsedirectsolver.analysis import compute_matching, matching_pairs, symbolic_analysis


def _rng(s):
    return np.random.default_rng(s)


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return M.csr_arrays(A)


def _absdense(A):
    return np.abs(np.asarray(sp.csr_matrix(A).todense()))


# ------------------------------------------------------------------------------------------------
# MC64

UNSYM = [("unsym_40", lambda: random_unsym(40, 0.12, _rng(1), diag=0.0)),
         ("unsym_60_scaled", lambda: sp.diags(np.exp(3 * _rng(2).standard_normal(60))) @ random_unsym(60, 0.08, _rng(3))),
         ("unsym_25_dense", lambda: random_unsym(25, 0.5, _rng(4)))]


def _matched_vals(Ad, perm):
    return Ad[np.arange(Ad.shape[0]), perm - 1]


@pytest.mark.parametrize("name,make", UNSYM, ids=[u[0] for u in UNSYM])
def test_job5_maximizes_the_product_with_dual_scalings(name, make):
    A = make()
    n = A.shape[0]
    rp, cv, nz = _csr(A)
    Ad = _absdense(A)
    m = compute_matching(rp, cv, nz, n, "G", "F", "O", 5)
    assert sorted(m.perm) == list(range(1, n + 1))
    srank = int((maximum_bipartite_matching(sp.csr_matrix(Ad > 0), perm_type="column") >= 0).sum())
    assert int(m.matched.sum()) == srank
    if srank < n:
        return
    # the optimum of sum log|a| over perfect matchings (scipy's assignment solver)
    with np.errstate(divide="ignore"):
        cost = np.where(Ad > 0, -np.log(np.where(Ad > 0, Ad, 1.0)), 1e30)
    r, c = linear_sum_assignment(cost)
    best = cost[r, c].sum()
    got = -np.log(_matched_vals(Ad, m.perm)).sum()
    assert abs(got - best) <= 1e-9 * max(1.0, abs(best))
    S = m.rscale[:, None] * Ad * m.cscale[None, :]
    assert S.max() <= 1 + 1e-12
    assert np.allclose(_matched_vals(S, m.perm), 1.0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("name,make", UNSYM, ids=[u[0] for u in UNSYM])
def test_job4_maximizes_the_sum(name, make):
    A = make()
    n = A.shape[0]
    rp, cv, nz = _csr(A)
    Ad = _absdense(A)
    m = compute_matching(rp, cv, nz, n, "G", "F", "O", 4)
    if int(m.matched.sum()) < n:
        return
    cost = np.where(Ad > 0, -Ad, 1e30)
    r, c = linear_sum_assignment(cost)
    assert abs(_matched_vals(Ad, m.perm).sum() - Ad[r, c].sum()) <= 1e-10 * Ad[r, c].sum()
    assert np.all(m.rscale == 1) and np.all(m.cscale == 1)


@pytest.mark.parametrize("job", [2, 3])
@pytest.mark.parametrize("name,make", UNSYM, ids=[u[0] for u in UNSYM])
def test_jobs_2_3_maximize_the_smallest_matched_entry(name, make, job):
    A = make()
    n = A.shape[0]
    rp, cv, nz = _csr(A)
    Ad = _absdense(A)
    m = compute_matching(rp, cv, nz, n, "G", "F", "O", job)
    full = int((maximum_bipartite_matching(sp.csr_matrix(Ad > 0), perm_type="column") >= 0).sum())
    assert int(m.matched.sum()) == full
    if full < n:
        return
    # the bottleneck by brute force: the largest threshold with a perfect matching among |a| >= t
    vals = np.unique(Ad[Ad > 0])
    best = max(t for t in vals
               if (maximum_bipartite_matching(sp.csr_matrix(Ad >= t), perm_type="column") >= 0).sum() == n)
    assert _matched_vals(Ad, m.perm).min() == best


def test_job1_structural_rank_and_completion():
    A = random_unsym(40, 0.05, _rng(7), diag=0.0)
    A = A.tolil()
    A[:, 3] = 0
    A[:, 17] = 0  # two empty columns: structurally singular
    A = A.tocsr()
    A.eliminate_zeros()
    rp, cv, nz = _csr(A)
    m = compute_matching(rp, cv, nz, 40, "G", "F", "O", 1)
    srank = int((maximum_bipartite_matching(sp.csr_matrix(A != 0), perm_type="column") >= 0).sum())
    assert srank <= 38 and int(m.matched.sum()) == srank
    assert sorted(m.perm) == list(range(1, 41))
    # matched rows sit on stored entries (job 1 counts every stored entry)
    pat = sp.csr_matrix(A != 0).todense()
    rows = np.nonzero(m.matched)[0]
    assert np.all(np.asarray(pat)[rows, m.perm[rows] - 1])


def test_symmetric_scaling_bounds_the_scaled_matrix():
    K = kkt_matrix(40, 20, 1e-8, _rng(9))
    L = M.lower_csr(K)
    rp, cv, nz = M.csr_arrays(L)
    m = compute_matching(rp, cv, nz, K.shape[0], "S", "L", "O", 5)
    assert m.symmetric and np.array_equal(m.rscale, m.cscale)
    Kd = _absdense(K)
    S = m.rscale[:, None] * Kd * m.rscale[None, :]
    assert S.max() <= 1 + 1e-12
    # the cycles' pairs: disjoint, distinct, on entries of the matrix
    prs = matching_pairs(m, rp, cv, nz, K.shape[0], "S", "L", "O", "all")
    flat = prs.ravel()
    assert len(set(flat.tolist())) == flat.size
    assert np.all(Kd[prs[:, 0] - 1, prs[:, 1] - 1] > 0)


# ------------------------------------------------------------------------------------------------
# pivot pairs

def _structural_zero_pivots_bruteforce(A, perm0, cand, pairs0):
    """Brute-force restatement: a candidate (not paired) at step k has a structurally zero pivot
    when the leading block (candidate diagonals absent, except those fixed by an earlier report;
    a pair a, c with c right after a counted as a 2x2 block) has a smaller structural rank than
    its size. perm0, pairs0 0-based."""
    n = A.shape[0]
    P = (sp.csr_matrix(A) != 0).tolil()
    for i in range(n):
        if cand[i]:
            P[i, i] = 0
    mate = {}
    for a, b in pairs0:
        mate[a], mate[b] = b, a
    out = []
    pos = np.empty(n, dtype=int)
    pos[perm0] = np.arange(n)
    fixed = set()
    for k in range(n):
        v = perm0[k]
        m = mate.get(v)
        if m is not None and pos[m] == k - 1:
            continue
        if not cand[v] or m is not None:
            continue
        idx = perm0[:k + 1]
        B = P[idx][:, idx].tocsr().astype(bool)
        B = B.tolil()
        for f in fixed:
            if pos[f] <= k:
                B[pos[f], pos[f]] = True
        sr = int((maximum_bipartite_matching(sp.csr_matrix(B), perm_type="column") >= 0).sum())
        if sr < k + 1:
            out.append(v)
            fixed.add(v)
    return out


def _candidates(A, tol=1e-6):
    Ad = _absdense(A)
    off = Ad - np.diag(np.diag(Ad))
    amax = off.max(axis=1)
    d = np.abs(np.diag(np.asarray(sp.csr_matrix(A).todense())))
    present = np.array([sp.csr_matrix(A)[i, i] != 0 or i in set(sp.csr_matrix(A)[i].indices) for i in range(A.shape[0])])
    return (~present) | (d <= tol * amax)


KKTS = [("kkt_60_30_delta0", lambda: kkt_matrix(60, 30, 0.0, _rng(3))),
        ("kkt_80_50_delta0", lambda: kkt_matrix(80, 50, 0.0, _rng(5))),
        ("symrand_70_zero_diag", lambda: random_symmetric(70, 0.05, _rng(6), 0.0))]


@pytest.mark.parametrize("mode", ["default", "all"])
@pytest.mark.parametrize("name,make", KKTS, ids=[k[0] for k in KKTS])
def test_pair_analysis_invariants(name, make, mode):
    """Every pair: partner right before its candidate in the ordering, both columns in one
    supernode and with the union structure below; the candidates of "all" are the rows whose
    diagonal is absent or negligible; with "default" no structurally zero pivot is left
    unpaired when a partner exists (brute force over the elimination prefixes)."""
    A = make()
    n = A.shape[0]
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    S = symbolic_analysis(rp, cv, n, "S", "L", values=nz, pivot_pairs=mode, amalgamation=(32, 0.25, 8))
    prs = S.pivot_pairs - 1
    assert prs.shape[0] > 0
    perm = S.etree_perm - 1
    pos = np.empty(n, dtype=int)
    pos[perm] = np.arange(n)
    fpos = np.empty(n, dtype=int)
    fpos[S.perm - 1] = np.arange(n)
    col2sn = S.col2sn
    cand = _candidates(A)
    for a, c in prs:
        assert pos[c] == pos[a] + 1, "the partner is not right before its candidate"
        assert col2sn[fpos[a]] == col2sn[fpos[c]], "a pair is split across supernodes"
        assert cand[c]
        if mode == "all":
            assert not cand[a]
    if mode == "all":
        paired = set(prs[:, 1].tolist())
        Ad = _absdense(A)
        for i in np.nonzero(cand)[0]:
            if i not in paired:  # an unpaired candidate has no free non-candidate neighbour left
                assert all(cand[j] or j in set(prs[:, 0].tolist()) for j in np.nonzero(Ad[i])[0] if j != i)
    else:
        # the search stops when a round adds no pair: a structurally zero pivot left over has no
        # free partner with |a_ij| >= u max_k |a_ik| (u = pivot_threshold, 0.01)
        left = _structural_zero_pivots_bruteforce(A, perm, cand, [tuple(p) for p in prs])
        Ad = _absdense(A)
        inpair = set(prs.ravel().tolist())
        for i in left:
            off = Ad[i].copy()
            off[i] = 0
            free = [j for j in np.nonzero(off)[0] if j not in inpair and off[j] >= 0.01 * off.max()]
            assert free == [] or S.pair_rounds == 8, f"zero pivot {i} left with free partners {free}"


def test_structural_zero_pivots_match_bruteforce_without_pairs():
    """The structurally zero pivots of the plain ordering (what "default" pairs in its first
    round) against the brute-force recurrence: the pairs of the first round are exactly the
    candidates it reports that have a partner."""
    A = kkt_matrix(50, 30, 0.0, _rng(11))
    n = A.shape[0]
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    S0 = symbolic_analysis(rp, cv, n, "S", "L", values=nz, pivot_pairs="none", amalgamation=(32, 0.25, 8))
    perm0 = S0.etree_perm - 1
    cand = _candidates(A)
    zs = _structural_zero_pivots_bruteforce(A, perm0, cand, [])
    assert len(zs) > 0
    S1 = symbolic_analysis(rp, cv, n, "S", "L", values=nz, pivot_pairs="default", amalgamation=(32, 0.25, 8))
    # the first round pairs only reported zero pivots (a candidate may also be taken as a partner)
    firsts = set(S1.pivot_pairs[:, 1].tolist() + S1.pivot_pairs[:, 0].tolist())
    assert set((np.array(zs) + 1).tolist()) <= firsts | set()


def test_no_pairs_gives_the_plain_analysis():
    """An SPD-like symmetric matrix has no candidates: "default" and "all" give the analysis of
    "none" array for array."""
    A = random_symmetric(60, 0.05, _rng(12), 5.0)
    A = (A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1)).tocsr()
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    ref = symbolic_analysis(rp, cv, A.shape[0], "S", "L", values=nz, pivot_pairs="none")
    for mode in ("default", "all"):
        S = symbolic_analysis(rp, cv, A.shape[0], "S", "L", values=nz, pivot_pairs=mode)
        assert S.pivot_pairs.shape[0] == 0
        for k in ("perm", "super_ptr", "rowval", "amap"):
            assert np.array_equal(S.maps[k], ref.maps[k])


def test_mutations_are_caught():
    """The brute-force recurrence distinguishes a wrong zero-pivot report: dropping the 'fixed by
    a partner' rule (a reported candidate keeps its absent diagonal) reports more zeros on a KKT
    with delta = 0, so the checks above would fail for an implementation with that slip."""
    A = kkt_matrix(40, 25, 0.0, _rng(13))
    n = A.shape[0]
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    S0 = symbolic_analysis(rp, cv, n, "S", "L", values=nz, pivot_pairs="none", amalgamation=(32, 0.25, 8))
    perm0 = S0.etree_perm - 1
    cand = _candidates(A)
    good = _structural_zero_pivots_bruteforce(A, perm0, cand, [])

    def slipped():
        P = (sp.csr_matrix(A) != 0).tolil()
        for i in range(n):
            if cand[i]:
                P[i, i] = 0
        out = []
        for k in range(n):
            v = perm0[k]
            if not cand[v]:
                continue
            idx = perm0[:k + 1]
            B = sp.csr_matrix(P[idx][:, idx].tocsr().astype(bool))
            if int((maximum_bipartite_matching(B, perm_type="column") >= 0).sum()) < k + 1:
                out.append(v)
        return out

    assert len(slipped()) > len(good)


# ------------------------------------------------------------------------------------------------
# nested dissection and the automatic ordering choice (src/symbolic/ordering.jl:86-326)

def _levels(S):
    """Depth of the supernodal tree of an analysis (its schedule depth)."""
    par = S.snparent.astype(np.int64) - 1
    h = np.ones(S.ns, dtype=np.int64)
    for s in range(S.ns):
        if par[s] >= 0:
            h[par[s]] = max(h[par[s]], h[s] + 1)
    return int(h.max())


@pytest.mark.parametrize("name,make", [("lap3d_20", lambda: M.laplacian3d(20, 20, 20)),
                                       ("lap2d_60x50", lambda: M.laplacian2d(60, 50)),
                                       ("kkt_2000_1000", lambda: M.kkt_matrix(2000, 1000)),
                                       ("two_blocks", lambda: sp.block_diag([M.laplacian3d(9, 9, 9),
                                                                             M.laplacian2d(30, 30)]).tocsr())],
                         ids=["lap3d_20", "lap2d_60x50", "kkt_2000_1000", "two_blocks"])
def test_nested_dissection_and_the_automatic_choice(name, make):
    """Nested dissection ("algo4") is a permutation whose analysis passes the structural checks;
    on 3-D grids it needs fewer flops than minimum degree; "default" takes the ordering of lower
    cost = flops + 1e8 x schedule depth (ties: minimum degree)."""
    A = make()
    n = A.shape[0]
    st = "S" if name.startswith("kkt") else "SPD"
    L = M.lower_csr(A)
    rp, cv, nz = M.csr_arrays(L)
    res = {}
    for o in ("algo3", "algo4", "default"):
        S = symbolic_analysis(rp, cv, n, st, "L", ordering=o, amalgamation=(32, 0.25, 8), values=nz,
                              pivot_pairs="none")
        assert sorted(S.etree_perm.tolist()) == list(range(1, n + 1))
        res[o] = (S.flops_true + 1e8 * _levels(S), S)
    if name.startswith("lap3d"):
        assert res["algo4"][1].flops_true < res["algo3"][1].flops_true
    best = "algo4" if res["algo4"][0] < res["algo3"][0] else "algo3"
    assert np.array_equal(res["default"][1].etree_perm, res[best][1].etree_perm)
