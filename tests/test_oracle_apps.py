"""Pins for the oracle's further apps (oracle/gg_oracle_apps.c): BC.C
against textbook Brandes accumulation, BellmanFord.C against Dijkstra."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

import oracle

INF = 2147483647 // 2


def brandes_single_source(off, ids, root):
    """Textbook single-source Brandes: BFS path counts sigma, then
    dependencies delta accumulated from the deepest level up."""
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
    for w in reversed(order):
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
    # diamond 0->1,0->2,1->3,2->3 plus a tail 3->4: two shortest paths to 3
    off, ids = oracle.csr_from_edges(5, [0, 0, 1, 2, 3], [1, 2, 3, 3, 4])
    score, level, sigma, _ = oracle.bc(off, ids, 0)
    assert list(level) == [0, 1, 1, 2, 3]
    assert list(sigma) == [1, 1, 1, 2, 2]
    # delta(3) = 1 (vertex 4); delta(1) = delta(2) = 1/2 * (1 + 1) = 1; delta(0) = 1+1 + ... = 4
    assert np.allclose(score, [4.0, 1.0, 1.0, 1.0, 0.0])


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
