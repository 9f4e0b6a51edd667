"""GPU parity of the generic edgeMap operators behind the further apps
(BC.C's path sums, PageRankDelta.C's signed pushes) and of GA.transpose().

One edgeMap step is compared with a direct numpy evaluation of the app's
update rule on the same seeded inputs, on every branch the reference's
threshold rule picks (the branch itself must match the oracle's choice).
Bars: output sets exact; values <= 1e-12 rel (FP64) / 1e-5 (FP32) per
element, since the branches sum in different orders.
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle
import refcheck as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _path(off, ids, dense, sym=False):
    """The branch the reference's edgeMap takes for this frontier (oracle)."""
    og = oracle.Graph(off, ids, symmetric=sym, P=64)
    n = og.n
    a = np.zeros(n)
    b = np.zeros(n)
    out = np.zeros(n, np.uint8)
    dm, oe, path = oracle.C.c_int64(), oracle.C.c_int64(), oracle.C.c_int()
    rc = oracle.lib().ggo_edgemap_once(og.h, oracle.OP_PR, np.ascontiguousarray(dense, np.uint8).ctypes.data,
                                       None, 0, 0, -1, a.ctypes.data, b.ctypes.data, None, None, out,
                                       oracle.C.byref(dm), oracle.C.byref(oe), oracle.C.byref(path))
    assert rc == 0
    return path.value


def _edges(off, ids):
    n = len(off) - 1
    return np.repeat(np.arange(n), np.diff(off)), np.asarray(ids, np.int64)


def _close(got, want, tol):
    scale = np.maximum(np.abs(want), 1e-300)
    return bool(np.all(np.abs(got - want) <= tol * scale + 1e-300))


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("frac", [0.0005, 0.02, 0.3])
def test_edgemap_pathsum_each_path(ctx, frac, dt, tol):
    """BC_F (app/BC.C:28-77): for d not visited, a[d] += sum of a[s] over
    frontier in-neighbours s; d is output when a[d] was 0 before."""
    print(f"pathsum frac={frac} {np.dtype(dt).name}", flush=True)
    off, ids = oracle.rmat_csr(14, 16, 5)
    n = len(off) - 1
    rng = np.random.default_rng(int(frac * 1e4) + 3)
    front = (rng.random(n) < frac).astype(np.uint8)
    vis = ((rng.random(n) < 0.3) | front.astype(bool)).astype(np.uint8)
    a0 = np.where(rng.random(n) < 0.5, 0.0, rng.integers(1, 50, n)).astype(dt)
    a0[front.astype(bool)] = rng.integers(1, 50, int(front.sum())).astype(dt)
    src, dst = _edges(off, ids)
    use = front[src].astype(bool) & ~vis[dst].astype(bool)
    acc = np.zeros(n)
    np.add.at(acc, dst[use], a0[src[use]].astype(np.float64))
    hit = np.zeros(n, bool)
    hit[dst[use]] = True
    want = a0.astype(np.float64) + acc
    want_out = hit & (a0 == 0)
    g = gg.Graph.from_csr(ctx, off, ids).partition(64)
    F = gg.Frontier.from_dense(g, front)
    A = gg.DeviceArray.from_host(ctx, a0)
    V = gg.DeviceArray.from_host(ctx, vis)
    out, gpath = gg.edge_map(g, F, gg.OP_PATHSUM, A, V)
    assert gpath == _path(off, ids, front)
    assert np.array_equal(out.dense().astype(bool), want_out)
    assert _close(A.to_host().astype(np.float64), want, tol)
    g.close()


def test_edgemap_pathsum_zero_and_negative_values(ctx):
    """A frontier value of 0 still makes d an output (its old value was 0),
    and negative sums are applied: the update is a plain add."""
    off, ids = oracle.csr_from_edges(5, [0, 1, 2, 3], [4, 4, 3, 2])
    a0 = np.array([0.0, 0.0, -2.0, 0.0, 0.0])
    vis = np.array([1, 1, 1, 0, 0], np.uint8)
    g = gg.Graph.from_csr(ctx, off, ids)
    F = gg.Frontier.from_dense(g, np.array([1, 1, 1, 0, 0], np.uint8))
    A = gg.DeviceArray.from_host(ctx, a0)
    V = gg.DeviceArray.from_host(ctx, vis)
    out, _ = gg.edge_map(g, F, gg.OP_PATHSUM, A, V, threshold=0)  # dense branch
    assert list(out.dense()) == [0, 0, 0, 1, 1]
    assert list(A.to_host()) == [0.0, 0.0, -2.0, -2.0, 0.0]
    g.close()


@pytest.mark.parametrize("frac", [0.0005, 0.02, 0.3])
def test_edgemap_pagerank_signed_changes(ctx, frac):
    """PR_F of app/PageRankDelta.C:27-72 pushes signed changes:
    nghSum[d] += Delta[s]/outdeg(s) over frontier sources, and every d with
    a frontier in-neighbour is output, whatever the sign of its sum."""
    print(f"signed pagerank frac={frac}", flush=True)
    off, ids = oracle.rmat_csr(14, 16, 6)
    n = len(off) - 1
    rng = np.random.default_rng(int(frac * 1e4) + 9)
    front = (rng.random(n) < frac).astype(np.uint8)
    delta = rng.standard_normal(n) * 1e-4
    s0 = rng.standard_normal(n) * 1e-5
    deg = np.diff(off)
    src, dst = _edges(off, ids)
    use = front[src].astype(bool)
    acc = np.zeros(n)
    np.add.at(acc, dst[use], delta[src[use]] / deg[src[use]])
    hit = np.zeros(n, bool)
    hit[dst[use]] = True
    want = s0 + acc
    g = gg.Graph.from_csr(ctx, off, ids).partition(64)
    F = gg.Frontier.from_dense(g, front)
    A = gg.DeviceArray.from_host(ctx, delta)
    B = gg.DeviceArray.from_host(ctx, s0)
    out, gpath = gg.edge_map(g, F, gg.OP_PAGERANK, A, B)
    assert gpath == _path(off, ids, front)
    assert np.array_equal(out.dense().astype(bool), hit)
    got = B.to_host()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    g.close()


@pytest.mark.parametrize("kind", ["spmv", "sum"])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("frac", [0.0005, 0.02, 0.3])
def test_edgemap_weighted_sums_partial_frontier(ctx, frac, dt, tol, kind):
    """SPMV_F (app/SPMV.C:7-48): p_next[d] += w * p_curr[s]; PR_F of
    app/PageRankConverage.C:34-88: p_next[d] += add_factor[s], weights not
    used.  Over a partial frontier, in FP64 and FP32, on every branch."""
    print(f"{kind} frac={frac} {np.dtype(dt).name}", flush=True)
    off, ids = oracle.rmat_csr(14, 16, 8)
    n = len(off) - 1
    w = oracle.weights(8, len(ids))
    rng = np.random.default_rng(int(frac * 1e4) + 21)
    front = (rng.random(n) < frac).astype(np.uint8)
    x = rng.standard_normal(n).astype(dt)
    y0 = rng.standard_normal(n).astype(dt)
    src, dst = _edges(off, ids)
    use = front[src].astype(bool)
    ww = w.astype(np.float64) if kind == "spmv" else np.ones(len(ids))
    acc = np.zeros(n)
    np.add.at(acc, dst[use], x[src[use]].astype(np.float64) * ww[use])
    hit = np.zeros(n, bool)
    hit[dst[use]] = True
    want = y0.astype(np.float64) + acc
    g = gg.Graph.from_csr(ctx, off, ids, w=w).partition(64)
    F = gg.Frontier.from_dense(g, front)
    X = gg.DeviceArray.from_host(ctx, x)
    Y = gg.DeviceArray.from_host(ctx, y0)
    out, gpath = gg.edge_map(g, F, gg.OP_SPMV if kind == "spmv" else gg.OP_SUM, X, Y)
    assert gpath == _path(off, ids, front)
    assert np.array_equal(out.dense().astype(bool), hit)
    got = Y.to_host().astype(np.float64)
    assert np.abs(got - want).max() <= tol * max(np.abs(want).max(), 1.0)
    g.close()


def _transpose_np(off, ids, w=None):
    n = len(off) - 1
    src, dst = _edges(off, ids)
    order = np.lexsort((src, dst))
    toff = np.zeros(n + 1, np.int64)
    np.add.at(toff, dst + 1, 1)
    toff = np.cumsum(toff)
    return toff.astype(np.uint32), src[order].astype(np.uint32), (None if w is None else w[order])


def test_transpose_swaps_directions(ctx):
    """GA.transpose() (graph-numa.h:547-555): BFS, PageRank and SpMV then
    run on the reversed edges; a second transpose restores the graph."""
    off, ids = oracle.rmat_csr(12, 16, 4)
    w = oracle.weights(4, len(ids))
    toff, tids, tw = _transpose_np(off, ids, w)
    g = gg.Graph.from_csr(ctx, off, ids, w=w).partition(64)
    g.transpose()
    o2, i2, w2 = g.csr()
    assert np.array_equal(o2, toff)
    for v in range(len(off) - 1):  # rows hold the same sources with their weights
        a, b = slice(toff[v], toff[v + 1]), slice(toff[v], toff[v + 1])
        ka, kb = np.argsort(i2[a], kind="stable"), np.argsort(tids[b], kind="stable")
        assert np.array_equal(i2[a][ka], tids[b][kb]) and np.array_equal(w2[a][ka], tw[b][kb])
    _, lv, _ = gg.bfs(g, 7)
    assert np.array_equal(lv.to_host(), R.bfs_levels(toff, tids, 7))
    want, it, _ = oracle.Graph(toff, tids, P=64).pagerank(10)
    got, info = gg.pagerank(g, np.float64)
    assert info.iters == it and R.rel_l1(got.to_host(), want) < 1e-12
    x = np.random.default_rng(1).random(len(off) - 1)
    y = gg.spmv(g, gg.DeviceArray.from_host(ctx, x))[0].to_host()
    assert R.rel_l1(y, R.to_scipy(off, ids, w) @ x) < 1e-12  # A x, not A^T x
    g.transpose()
    o3, i3, w3 = g.csr()
    assert np.array_equal(o3, off) and np.array_equal(i3, ids) and np.array_equal(w3, w)
    _, lv, _ = gg.bfs(g, 7)
    assert np.array_equal(lv.to_host(), R.bfs_levels(off, ids, 7))
    g.close()


def test_transpose_symmetric_is_identity(ctx):
    off, ids = oracle.rmat_csr(10, 8, 2, symmetric=True)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).partition(16)
    g.transpose()
    o, i, _ = g.csr()
    assert np.array_equal(o, off) and np.array_equal(i, ids)
    g.close()
