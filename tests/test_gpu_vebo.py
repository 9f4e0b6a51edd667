"""VEBO order (graphtools/VEBO.C) and graph relabelling on the GPU path.

The order must equal the oracle's bit for bit (oracle/gg_oracle_apps.c
ggo_vebo, pinned by a hand-worked case and the order's invariants).  The
relabelled graph must equal a numpy renumbering of the CSR, and the apps
must be equivariant under it: BFS levels exactly, PageRank ranks within
1e-12 rel L1 (only the summation order changes).
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


def _relabel_np(off, ids, new, w=None):
    n = len(off) - 1
    old_of_new = np.empty(n, np.int64)
    old_of_new[new] = np.arange(n)
    deg = np.diff(off.astype(np.int64))[old_of_new]
    noff = np.zeros(n + 1, np.int64)
    noff[1:] = np.cumsum(deg)
    nids = np.empty(len(ids), np.uint32)
    nw = None if w is None else np.empty(len(ids), np.float32)
    for i in range(n):
        o = old_of_new[i]
        seg = slice(off[o], off[o + 1])
        nids[noff[i]:noff[i + 1]] = new[ids[seg]]
        if w is not None:
            nw[noff[i]:noff[i + 1]] = w[seg]
    return noff.astype(np.uint32), nids, nw


@pytest.mark.parametrize("scale,sym,P", [(10, False, 16), (12, True, 64), (16, False, 384), (14, False, 7)])
def test_vebo_order_matches_oracle(ctx, scale, sym, P):
    print(f"vebo scale={scale} sym={sym} P={P}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 3, symmetric=sym)
    n = len(off) - 1
    want, wst = oracle.vebo(np.bincount(ids, minlength=n), P)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=sym)
    got, st = g.vebo_order(P)
    assert np.array_equal(got, want) and np.array_equal(st, wst)
    g.close()


def test_vebo_order_without_zero_degree_vertices(ctx):
    """A directed ring: every in-degree is 1, no zero-degree vertex to even
    out the vertex counts (the reference's fill rule would overrun)."""
    n = 1000
    off, ids = oracle.csr_from_edges(n, list(range(n)), [(v + 1) % n for v in range(n)])
    want, wst = oracle.vebo(np.ones(n, np.int64), 13)
    g = gg.Graph.from_csr(ctx, off, ids)
    got, st = g.vebo_order(13)
    assert np.array_equal(got, want) and np.array_equal(st, wst)
    with pytest.raises(gg.GGError):
        g.vebo_order(0)
    g.close()


def test_relabel_matches_numpy_and_apps_are_equivariant(ctx):
    off, ids = oracle.rmat_csr(12, 16, 5)
    n = len(off) - 1
    w = oracle.weights(5, len(ids))
    g = gg.Graph.from_csr(ctx, off, ids, w=w)
    new, _ = g.vebo_order(64)
    h = g.relabel(new)
    noff, nids, nw = _relabel_np(off, ids, new, w)
    o2, i2, w2 = h.csr()
    assert np.array_equal(o2, noff) and np.array_equal(i2, nids) and np.array_equal(w2, nw)
    root = 17
    _, lv, _ = gg.bfs(g, root)
    _, lv2, _ = gg.bfs(h, int(new[root]))
    assert np.array_equal(lv2.to_host()[new], lv.to_host())
    g.partition(64)
    h.partition(64)
    p1, i1 = gg.pagerank(g, np.float64)
    p2, i2_ = gg.pagerank(h, np.float64)
    assert i1.iters == i2_.iters
    assert R.rel_l1(p2.to_host()[new], p1.to_host()) < 1e-12
    with pytest.raises((gg.GGError, ValueError)):
        g.relabel(np.zeros(n, np.uint32))
    g.close()
    h.close()


def test_vebo_tool(tmp_path):
    """apps/vebo (graphtools/VEBO.C's command line): a Galois .gr in, the
    VEBO-renumbered graph out as PBBS AdjacencyGraph text."""
    import os
    import subprocess
    from test_frontend import write_galois
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "apps", "_build", "vebo")
    if not os.path.exists(exe):
        pytest.skip("apps not built")
    off, ids = oracle.rmat_csr(11, 8, 6)
    n = len(off) - 1
    src = str(tmp_path / "g.gr")
    dst = str(tmp_path / "g.adj")
    write_galois(src, off, ids)
    r = subprocess.run([exe, "-p", "32", "-r", "5", src, dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    new, _ = oracle.vebo(np.bincount(ids, minlength=n), 32)
    assert f"new: {new[5]} old: 5" in r.stderr
    tok = open(dst).read().split()
    assert tok[0] == "AdjacencyGraph" and int(tok[1]) == n and int(tok[2]) == len(ids)
    got_off = np.array(tok[3:3 + n], np.int64)
    got_ids = np.array(tok[3 + n:], np.int64)
    noff, nids, _ = _relabel_np(off, ids, new)
    assert np.array_equal(got_off, noff[:-1]) and np.array_equal(got_ids, nids)
