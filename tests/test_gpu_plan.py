"""The hot-sources-first plan (gg_graph_reorder) across a graph's life:
partitioning, transposition (which drops it), re-planning, new weights and
option changes, with every app that uses it checked against the oracle
after each step.  Integer results must be exact, FP results within the
usual bars (<= 1e-12 rel L1 in FP64)."""
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


def _check_all(g, off, ids, w, root):
    """BFS levels and CC labels exact, SpMV and weighted PageRank within
    1e-12 of the oracle on the graph as it now stands."""
    n = len(off) - 1
    _, lv, _ = gg.bfs(g, root)
    assert np.array_equal(lv.to_host(), R.bfs_levels(off, ids, root))
    want_cc, _, _ = oracle.Graph(off, ids, P=16).cc()
    lab, _ = gg.components(g)
    assert np.array_equal(lab.to_host(), want_cc)
    x = oracle.uniform_pm1(2, n)
    y, _ = gg.spmv(g, gg.DeviceArray.from_host(g.ctx, x))
    assert R.rel_l1(y.to_host(), oracle.spmv_csr(off, ids, w, x)) < 1e-12
    want, it, _ = oracle.wpagerank(off, ids, w, 0.85, 1e-7, 100)
    r, info = gg.wpagerank(g, np.float64, 0.85, 1e-7, 100)
    assert info.iters == it and R.rel_l1(r.to_host(), want) < 1e-12


def test_plan_lifecycle(ctx):
    off, ids = oracle.rmat_csr(12, 8, 31)
    m = len(ids)
    w = oracle.weights(31, m)
    g = gg.Graph.from_csr(ctx, off, ids, w=w)
    g.reorder(1)
    _check_all(g, off, ids, w, 3)
    g.partition(32)  # partitions leave the plan in place
    _check_all(g, off, ids, w, 3)
    # transposition drops the plan; the transposed graph runs plain
    g.transpose()
    toff, tids, tw = g.csr()
    _check_all(g, toff, tids, tw, 3)
    g.reorder(1)  # a plan of the transposed graph
    _check_all(g, toff, tids, tw, 3)
    g.transpose()  # back, plan dropped again
    g.reorder(1)
    # integer lengths for Bellman-Ford: weights replaced, plan rebuilt
    wl = np.round(w * 16).astype(np.float32) + 1
    g.set_weights(wl)
    d, _ = gg.sssp(g, 3)
    want_d, _ = oracle.sssp(off, ids, wl, 3)
    assert np.array_equal(d.to_host().astype(np.int64), want_d)
    # row-block width changed: the plan's schedules follow
    g.set_option(gg.OPT_ROW_EDGES_PER_LANE, 4)
    _, lv, _ = gg.bfs(g, 3)
    assert np.array_equal(lv.to_host(), R.bfs_levels(off, ids, 3))
    x = oracle.uniform_pm1(5, len(off) - 1)
    gy = oracle.uniform_pm1(6, len(off) - 1)
    gx_w, gw_w = oracle.spmv_bwd(off, ids, wl, x, gy)
    gx, gw, _ = gg.spmv_bwd(g, gg.DeviceArray.from_host(ctx, x), gg.DeviceArray.from_host(ctx, gy))
    assert R.rel_l1(gx.to_host(), gx_w) < 1e-12
    assert np.array_equal(gw.to_host()[:m], gw_w)
    g.reorder(0)
    _, lv, _ = gg.bfs(g, 3)
    assert np.array_equal(lv.to_host(), R.bfs_levels(off, ids, 3))
    g.close()
