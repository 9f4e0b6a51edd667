"""The source-segmented pull (GG_OPT_XCD_SEGMENTS) of PageRank.C's loop and
of SpMV: each XCD gathers from one source group (its hottest sources cached
in LDS), every (row, group) pair's partial sum is written once and a combine
pass adds them in group order.  Against the oracle (PageRank.C and SpMV,
FP64 within 1e-12, FP32 within 1e-5), against the plain pull on the same
plan, bitwise equal run to run (which work-group and XCD serve a group
change no value), on R-MAT graphs, a hub whose in-edges from one group span
several work items (long rows split into chunks), and tiny graphs.
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle
import refcheck as R

pytestmark = pytest.mark.gpu

OPT_XCD = 3


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _run(g, dt, mode):
    g.set_option(OPT_XCD, mode)
    r, info = gg.pagerank(g, dt)
    return r.to_host(), info


@pytest.mark.parametrize("scale,P", [(10, 16), (14, 64), (18, 384), (20, 384)])
def test_segmented_pagerank_vs_oracle(ctx, scale, P):
    print(f"segmented pagerank scale={scale}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 1)
    want, it, _ = oracle.Graph(off, ids, P=P).pagerank(10)
    g = gg.Graph.from_csr(ctx, off, ids).partition(P)
    g.reorder(1)
    a, ia = _run(g, np.float64, 2)
    b, _ = _run(g, np.float64, 2)
    assert ia.iters == it
    assert np.array_equal(a, b), "segmented ranks differ between runs"
    assert R.rel_l1(a, want) < 1e-12
    plain, _ = _run(g, np.float64, 0)
    assert R.rel_l1(a, plain) < 1e-13
    f32, _ = _run(g, np.float32, 2)
    assert R.rel_l1(f32, want) < 1e-5
    g.close()


def test_segmented_pagerank_hub_and_chunks(ctx):
    """One destination with 40,000 in-edges (several 512-edge chunks in each
    source group) plus an R-MAT graph around it."""
    off, ids = oracle.rmat_csr(16, 8, 7)
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    hub_src = np.arange(1, 40001) % n
    s2 = np.concatenate([src, hub_src])
    d2 = np.concatenate([ids.astype(np.int64), np.full(len(hub_src), 5)])
    off2, ids2 = oracle.csr_from_edges(n, s2, d2)
    want, it, _ = oracle.Graph(off2, ids2, P=64).pagerank(10)
    g = gg.Graph.from_csr(ctx, off2, ids2).partition(64)
    g.reorder(1)
    a, info = _run(g, np.float64, 2)
    assert info.iters == it and R.rel_l1(a, want) < 1e-12
    b, _ = _run(g, np.float64, 2)
    assert np.array_equal(a, b)
    g.close()


# (a graph whose frontier runs out of out-edges is left out: PageRank.C's
# next edgeMap on the empty() frontier aborts in the reference)
@pytest.mark.parametrize("edges", [[(0, 1), (1, 2), (2, 0)], [(0, 1), (1, 0)], [(0, 2), (1, 2), (3, 2), (2, 0)]])
def test_segmented_pagerank_tiny(ctx, edges):
    s, d = zip(*edges)
    n = max(max(s), max(d)) + 1
    off, ids = oracle.csr_from_edges(n, s, d)
    want, it, _ = oracle.Graph(off, ids, P=1).pagerank(10)
    g = gg.Graph.from_csr(ctx, off, ids).partition(1)
    g.reorder(1)
    a, info = _run(g, np.float64, 2)
    assert info.iters == it and R.rel_l1(a, want) < 1e-12
    g.close()
    # the segmented SpMV on the same graph, weighted
    w = (np.arange(len(ids), dtype=np.float32) + 1) / 4
    x = np.arange(1, n + 1, dtype=np.float64)
    g = gg.Graph.from_csr(ctx, off, ids, w=w).partition(1)
    g.reorder(1)
    g.set_option(OPT_XCD, 2)
    y = gg.spmv(g, gg.DeviceArray.from_host(ctx, x))[0].to_host()
    assert R.rel_l1(y, oracle.spmv_csr(off, ids, w, x)) < 1e-12
    g.close()


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("scale", [12, 16])
def test_segmented_spmv_vs_oracle(ctx, scale, weighted):
    """SpMV over the segmented pull (each term x[s] times its edge's weight,
    the weights sorted into the segmented order beside the ids): FP64 within
    1e-12 and FP32 within 1e-5 of the FP64 oracle, bitwise equal run to run,
    and against the plain pull on the same plan."""
    print(f"segmented spmv scale={scale} weighted={weighted}", flush=True)
    g = gg.Graph.rmat(ctx, scale, 16, 3, weighted=weighted)
    off, ids, w = g.csr()
    x = np.random.default_rng(scale).uniform(-1, 1, g.n)
    g.partition(64)
    g.reorder(1)
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        xd = gg.DeviceArray.from_host(ctx, x.astype(dt))
        want = oracle.spmv_csr(off, ids, w if weighted else None, x.astype(dt).astype(np.float64))
        g.set_option(OPT_XCD, 2)
        a = gg.spmv(g, xd)[0].to_host()
        b = gg.spmv(g, xd)[0].to_host()
        assert np.array_equal(a, b), "segmented SpMV differs between runs"
        assert R.rel_l1(a, want) < tol
        g.set_option(OPT_XCD, 0)
        plain = gg.spmv(g, xd)[0].to_host()
        assert R.rel_l1(a, plain) < tol
    g.close()


def test_segmented_spmv_hub_and_chunks(ctx):
    """Weighted: one destination with 40,000 in-edges (several chunks in each
    source group) plus an R-MAT graph around it."""
    off, ids = oracle.rmat_csr(16, 8, 7)
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    hub_src = np.arange(1, 40001) % n
    s2 = np.concatenate([src, hub_src])
    d2 = np.concatenate([ids.astype(np.int64), np.full(len(hub_src), 5)])
    off2, ids2 = oracle.csr_from_edges(n, s2, d2)
    w = np.random.default_rng(8).uniform(0.0, 1.0, len(ids2)).astype(np.float32) + np.float32(1e-3)
    x = np.random.default_rng(9).uniform(-1, 1, n)
    want = oracle.spmv_csr(off2, ids2, w, x)
    g = gg.Graph.from_csr(ctx, off2, ids2, w=w).partition(64)
    g.reorder(1)
    g.set_option(OPT_XCD, 2)
    y = gg.spmv(g, gg.DeviceArray.from_host(ctx, x))[0].to_host()
    assert R.rel_l1(y, want) < 1e-12
    g.close()


def test_segmented_option_values(ctx):
    g = gg.Graph.rmat(ctx, 8, 16, 1)
    assert g.get_option(OPT_XCD) == 1
    for v in (0, 1, 2):
        g.set_option(OPT_XCD, v)
        assert g.get_option(OPT_XCD) == v
    with pytest.raises(gg.GGError):
        g.set_option(OPT_XCD, 3)
    # without a plan the option changes nothing (the plain pull runs)
    g.partition(16)
    g.set_option(OPT_XCD, 2)
    r, info = gg.pagerank(g, np.float64)
    off, ids, _ = g.csr()
    want, it, _ = oracle.Graph(off, ids, P=16).pagerank(10)
    assert info.iters == it and R.rel_l1(r.to_host(), want) < 1e-12
    g.close()
