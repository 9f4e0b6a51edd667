"""Pins for the oracle's further apps (oracle/gg_oracle_apps.c): BC.C
against textbook Brandes accumulation, BellmanFord.C against Dijkstra,
PageRankDelta.C against PageRank power iteration and a one-step closed form,
PageRankConverage.C against its first step in closed form and a direct solve,
graphtools/VEBO.C against a hand-worked order and its invariants."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

import oracle
import refcheck as R

INF = 2147483647 // 2


def brandes_single_source(off, ids, root, stored_counts=True):
    """Textbook single-source Brandes: BFS path counts sigma, then
    dependencies delta accumulated from the deepest level up.  With
    stored_counts, the reference's edgeMap rule on the Levels' stored
    out-degree sums (ligra-numa.h:946-947) is applied: a level whose
    vertices have no out-edges passes nothing up (only the deepest level of
    a directed graph can be such a level)."""
    n = len(off) - 1
    dist = np.full(n, -1, np.int64)
    sigma = np.zeros(n)
    dist[root] = 0
    sigma[root] = 1
    order = [root]
    q = [root]
    while q:
        nq = []
        for v in q:
            for w in ids[off[v]:off[v + 1]]:
                if dist[w] < 0:
                    dist[w] = dist[v] + 1
                    nq.append(w)
                    order.append(w)
                if dist[w] == dist[v] + 1:
                    sigma[w] += sigma[v]
        q = nq
    delta = np.zeros(n)
    # predecessors of w: v with an edge v->w and dist[v] = dist[w]-1
    preds = {}
    for v in range(n):
        for w in ids[off[v]:off[v + 1]]:
            if dist[v] >= 0 and dist[w] == dist[v] + 1:
                preds.setdefault(int(w), []).append(v)
    deepest = dist.max()
    dead = stored_counts and all(off[v + 1] == off[v] for v in np.flatnonzero(dist == deepest))
    for w in reversed(order):
        if dead and dist[w] == deepest:
            continue
        for v in preds.get(int(w), []):
            delta[v] += sigma[v] / sigma[w] * (1 + delta[w])
    return delta, dist, sigma


@pytest.mark.parametrize("scale,ef,seed,sym,root", [(6, 8, 2, True, 3), (8, 4, 2, True, 0), (7, 8, 5, False, 1),
                                                    (9, 16, 1, False, 100)])
def test_bc_matches_brandes(scale, ef, seed, sym, root):
    off, ids = oracle.rmat_csr(scale, ef, seed, symmetric=sym)
    score, level, sigma, rounds = oracle.bc(off, ids, root)
    delta, dist, sig = brandes_single_source(off, ids, root)
    reached = dist >= 0
    assert np.array_equal(level, np.where(reached, dist, -1))
    assert np.allclose(sigma[reached], sig[reached], rtol=1e-12)
    assert np.all(np.isnan(score[~reached]))
    assert np.allclose(score[reached], delta[reached], rtol=1e-10, atol=1e-10)
    assert rounds == dist.max() + 1


def test_bc_tiny_by_hand():
    # diamond 0->1,0->2,1->3,2->3 plus a tail 3->4 and an edge 4->0 back to
    # the root: two shortest paths to 3
    off, ids = oracle.csr_from_edges(5, [0, 0, 1, 2, 3, 4], [1, 2, 3, 3, 4, 0])
    score, level, sigma, _ = oracle.bc(off, ids, 0)
    assert list(level) == [0, 1, 1, 2, 3]
    assert list(sigma) == [1, 1, 1, 2, 2]
    # delta(3) = 1 (vertex 4); delta(1) = delta(2) = 1/2 * (1 + 1) = 1; delta(0) = 2 * (1 + 1) = 4
    assert np.allclose(score, [4.0, 1.0, 1.0, 1.0, 0.0])


def test_bc_deepest_level_without_out_edges():
    """The same diamond without 4->0: the deepest level {4} has no
    out-edges, so its stored out-degree sum is 0 and the reference's first
    backward edgeMap returns empty(): 3 gets no dependency from 4
    (textbook Brandes would give [4, 1, 1, 1, 0]).  The compiled reference
    gives [0, .5, .5, 0, 0] here: its root also has no in-edge, and after
    GA.transpose() its compressed CSC still lists only vertices with
    original in-edges, so the dense backward step to the root skips it.
    Here every structure follows the transpose, so the root keeps its
    dependencies: [3, .5, .5, 0, 0]."""
    off, ids = oracle.csr_from_edges(5, [0, 0, 1, 2, 3], [1, 2, 3, 3, 4])
    score, level, sigma, rounds = oracle.bc(off, ids, 0)
    assert list(level) == [0, 1, 1, 2, 3] and rounds == 4
    assert np.allclose(score, [3.0, 0.5, 0.5, 0.0, 0.0])
    textbook, _, _ = brandes_single_source(off, ids, 0, stored_counts=False)
    assert np.allclose(textbook, [4.0, 1.0, 1.0, 1.0, 0.0])


@pytest.mark.parametrize("scale,seed,sym", [(8, 3, False), (10, 4, True)])
def test_sssp_matches_dijkstra(scale, seed, sym):
    off, ids = oracle.rmat_csr(scale, 8, seed, symmetric=sym)
    w = (1 + (np.arange(len(ids)) * 2654435761 % 19)).astype(np.float32)
    dist, rounds = oracle.sssp(off, ids, w, 0)
    n = len(off) - 1
    A = sp.csr_matrix((w.astype(np.float64), ids.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    want = dijkstra(A, indices=0)
    want = np.where(np.isinf(want), INF, want).astype(np.int64)
    assert np.array_equal(dist, want)
    assert rounds >= 1


def test_sssp_negative_cycle():
    off, ids = oracle.csr_from_edges(3, [0, 1, 2], [1, 2, 1])
    w = np.array([1, 1, -3], np.float32)
    dist, rounds = oracle.sssp(off, ids, w, 0)
    assert rounds == 3 and np.all(dist == -INF)


def power_pagerank(off, ids, damping=0.85, eps=1e-7, max_iter=10000):
    """Textbook PageRank power iteration with scipy (no dangling-mass
    correction, as the reference apps): p_k = (1-d)/n + d A^T (p_{k-1}/deg)
    from the uniform vector, until ||p_k - p_{k-1}||_1 < eps.  (p, k)."""
    n = len(off) - 1
    deg = np.diff(off).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(off))
    AT = sp.csr_matrix((np.ones(len(ids)), (ids, src)), shape=(n, n))
    p = np.full(n, 1.0 / n)
    for k in range(1, max_iter + 1):
        q = (1 - damping) / n + damping * (AT @ np.where(deg > 0, p / np.maximum(deg, 1), 0.0))
        done = np.abs(q - p).sum() < eps
        p = q
        if done:
            return p, k
    return p, max_iter


@pytest.mark.parametrize("scale,seed", [(8, 1), (11, 4)])
def test_prdelta_without_filter_is_power_iteration(scale, seed):
    """With eps2 = 0 every vertex with a non-zero change stays active, and
    the delta recurrence p_k = p_{k-1} + Delta_k, Delta_k = d A^T
    (Delta_{k-1}/deg), is the power iteration exactly: same ranks (to
    rounding) and the same round at which the L1 change drops below eps."""
    off, ids = oracle.rmat_csr(scale, 8, seed)
    want, it = power_pagerank(off, ids)
    got, rounds = oracle.prdelta(off, ids, eps2=0.0)
    assert rounds == it
    assert np.abs(got - want).sum() / np.abs(want).sum() < 1e-12


def test_prdelta_huge_filter_is_one_power_step():
    """With eps2 so large that no vertex stays active, round 2 pushes nothing,
    so p is one damped step from the uniform vector:
    p = (1-d)/n + d * A^T (1/n / outdeg) (scipy), after exactly 2 rounds."""
    off, ids = oracle.rmat_csr(9, 8, 7)
    want, _ = power_pagerank(off, ids, max_iter=1)
    got, rounds = oracle.prdelta(off, ids, eps2=1e300)
    assert rounds == 2
    assert np.allclose(got, want, rtol=1e-13, atol=0)


def test_prdelta_filter_converges_near_pagerank():
    """With the app's eps2 = 0.01 the filtered run stays close to the fixed
    point: its L1 distance to converged power iteration is small."""
    off, ids = oracle.rmat_csr(11, 16, 2)
    want, _ = power_pagerank(off, ids, eps=1e-13)
    got, rounds = oracle.prdelta(off, ids)
    assert 2 < rounds < 200
    assert np.abs(got - want).sum() / np.abs(want).sum() < 5e-2


def _dangling_system(off, ids, damping=0.85):
    n = len(off) - 1
    deg = np.diff(off).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(off))
    AT = sp.csr_matrix((damping / deg[src], (ids, src)), shape=(n, n))  # d A^T D^-1
    return n, deg, AT


def test_prconv_first_iteration_closed_form():
    """One iteration: x = d A^T (p0/deg); x += (1 - sum x)/n; x /= sum x,
    evaluated with scipy from the uniform vector (max_iter = 1)."""
    off, ids = oracle.rmat_csr(10, 8, 3)
    n, deg, AT = _dangling_system(off, ids)
    x = AT @ np.full(n, 1.0 / n)
    x = x + (1 - x.sum()) / n
    x = x / x.sum()
    got, it = oracle.prconv(off, ids, max_iter=1)
    assert it == 1 and np.allclose(got, x, rtol=1e-13, atol=0)


@pytest.mark.parametrize("scale,seed", [(8, 1), (10, 5)])
def test_prconv_converges_to_dangling_pagerank(scale, seed):
    """The fixed point spreads the non-forwarded mass uniformly: p = d A^T
    D^-1 p + (1 - d sum_{deg>0} p)/n with sum p = 1.  Solved directly
    (dense LU); the run ends within the app's 100 iterations once the L1
    change is below 1e-7, so p agrees to about 1e-6."""
    off, ids = oracle.rmat_csr(scale, 8, seed)
    n, deg, AT = _dangling_system(off, ids)
    fwd = (deg > 0).astype(np.float64)
    M = np.eye(n) - AT.toarray() + (0.85 / n) * np.outer(np.ones(n), fwd)
    want = np.linalg.solve(M, np.full(n, 1.0 / n))
    got, it = oracle.prconv(off, ids)
    assert it < 100
    assert abs(got.sum() - 1) < 1e-12
    assert np.abs(got - want).sum() < 1e-5


def test_vebo_by_hand():
    """n = 6, P = 2, in-degrees [3, 0, 1, 2, 0, 1] (worked through
    graphtools/VEBO.C by hand): sorted order 0, 3, 2, 5, 1, 4; phase 1 gives
    partition 0 the vertices of degree 3 and 1 (v0, v2), partition 1 those
    of degree 2 and 1 (v3, v5); the two zero-degree vertices fill each
    partition to n/P = 3; ids go group by group: p0 = [v0, v2, v1],
    p1 = [v3, v5, v4]."""
    new, starts = oracle.vebo(np.array([3, 0, 1, 2, 0, 1]), 2)
    assert list(new) == [0, 2, 1, 3, 5, 4]
    assert list(starts) == [0, 3, 6]


@pytest.mark.parametrize("scale,sym,P", [(10, False, 16), (12, True, 64), (14, False, 384)])
def test_vebo_invariants(scale, sym, P):
    off, ids = oracle.rmat_csr(scale, 16, 3, symmetric=sym)
    n = len(off) - 1
    indeg = np.bincount(ids, minlength=n).astype(np.int64)
    new, starts = oracle.vebo(indeg, P)
    assert np.array_equal(np.sort(new), np.arange(n))  # a permutation
    old_of_new = np.empty(n, np.int64)
    old_of_new[new] = np.arange(n)
    d_new = indeg[old_of_new]
    st = starts.astype(np.int64)
    assert st[0] == 0 and st[-1] == n and np.all(np.diff(st) >= 0)
    e = np.array([d_new[st[j]:st[j + 1]].sum() for j in range(P)])
    assert e.max() - e.min() <= indeg.max()  # edge-balanced to within one vertex
    for j in range(P):
        seg = d_new[st[j]:st[j + 1]]
        # degree groups high to low, ties by old id: (-degree, id) increases
        key = -seg * (n + 1) + old_of_new[st[j]:st[j + 1]]
        assert np.all(np.diff(key) > 0)
    zeros = int((indeg == 0).sum())
    if zeros >= P:
        sz = np.diff(st)
        assert sz.max() - sz.min() <= 1 or sz.max() <= n // P + 1


# ---- against the reference engine (tests/golden/apps_*.npz, written by
# tests/golden/make_golden.py from oracle/_ref, the reference compiled) ----

def _ids(p):
    return p.split("/")[-1]


@pytest.mark.parametrize("path", R.app_golden_cases("bc"), ids=_ids)
def test_bc_matches_reference(path):
    """The reference computes in FP32 (fType float); the oracle in FP64."""
    z = R.load_golden(path)
    want = z["bc"]
    got, lv, _, rounds = oracle.bc(z["off"], z["ids"], int(z["root"]))
    assert rounds == int(z["bc_iters"])
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.abs(got[ok] - want[ok]).sum() <= 1e-6 * max(np.abs(want[ok]).sum(), 1e-30)


@pytest.mark.parametrize("path", R.app_golden_cases("bellmanford"), ids=_ids)
def test_bellman_ford_matches_reference(path):
    z = R.load_golden(path)
    got, _ = oracle.sssp(z["off"], z["ids"], z["bellmanford_w"].astype(np.float32), int(z["root"]))
    assert np.array_equal(got, z["bellmanford"].astype(np.int64))


@pytest.mark.parametrize("path", R.app_golden_cases("prdelta"), ids=_ids)
def test_prdelta_matches_reference(path):
    z = R.load_golden(path)
    got, rounds = oracle.prdelta(z["off"], z["ids"])
    assert rounds == int(z["prdelta_iters"])
    assert np.abs(got - z["prdelta"]).sum() <= 1e-12 * np.abs(z["prdelta"]).sum()


@pytest.mark.parametrize("path", R.app_golden_cases("prconv"), ids=_ids)
def test_prconv_matches_reference(path):
    z = R.load_golden(path)
    got, it = oracle.prconv(z["off"], z["ids"])
    assert it == int(z["prconv_iters"])
    assert np.abs(got - z["prconv"]).sum() <= 1e-12 * np.abs(z["prconv"]).sum()
