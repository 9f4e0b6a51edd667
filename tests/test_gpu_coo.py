"""The COO traversal of edgeMap's dense COO branch (GG_OPT_DENSE_COO = 1,
ligra-numa.h:803-866): each destination's edges in its COO partition's
Hilbert order, the order the reference's PART96 loop adds them in, with one
owner per destination (no atomics, a fixed summation order).

Checked against the oracle, not against the default pull: one edgeMap step
per operator against the FP64 sum over the frontier's in-edges (numpy), the
PageRank.C loop (generic edgeMap and the fused loop) and SPMV.C against the
oracle's COO run, BFS/CC/BC/Bellman-Ford whole against the oracle.  Results
are bitwise equal run to run.  Bars: output frontiers and integer results
exact; FP64 within 1e-12 and FP32 within 1e-5 relative L1.
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


def _graph(ctx, scale, seed, sym=False, weighted=False, P=16):
    off, ids = oracle.rmat_csr(scale, 16, seed, symmetric=sym)
    w = oracle.weights(seed, len(ids)) if weighted else None
    g = gg.Graph.from_csr(ctx, off, ids, w=w, symmetric=sym).partition(P)
    return g, off, ids, w


def _both(g, front, kind, make):
    """edge_map under the pull (option 0) and the COO traversal (option 1);
    'make' returns fresh (a, b) device arrays; returns per option the output
    flags, the path taken and the arrays' host values."""
    res = []
    for opt in (0, 1):
        g.set_option(gg.OPT_DENSE_COO, opt)
        a, b = make()
        F = gg.Frontier.from_dense(g, front) if front is not None else gg.Frontier.bits(g)
        out, path = gg.edge_map(g, F, kind, a, b)
        assert path == gg.PATH_COO, path
        res.append((out.dense(), a.to_host(), None if b is None else b.to_host(), out.info()))
    g.set_option(gg.OPT_DENSE_COO, 0)
    return res


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("frac", [None, 0.7])
@pytest.mark.parametrize("kind", ["pagerank", "spmv", "sum"])
def test_coo_fp_sums_vs_fp64(ctx, kind, frac, dt, tol):
    """One dense COO step: y[d] += sum over active in-neighbours s of
    x[s]/outdeg(s) (PR_F), x[s]*w (SPMV_F) or x[s] (PageRankConverage's
    PR_F), against that sum in FP64; the output frontier is the set of
    destinations with an active in-neighbour; a second run is bitwise equal."""
    g, off, ids, w = _graph(ctx, 13, 3, weighted=(kind == "spmv"))
    n = g.n
    rng = np.random.default_rng(5)
    front = None if frac is None else (rng.random(n) < frac).astype(np.uint8)
    x = rng.random(n).astype(dt)
    y0 = rng.random(n).astype(dt)
    k = {"pagerank": gg.OP_PAGERANK, "spmv": gg.OP_SPMV, "sum": gg.OP_SUM}[kind]
    g.set_option(gg.OPT_DENSE_COO, 1)
    got = []
    for _ in range(2):
        a, b = gg.DeviceArray.from_host(ctx, x), gg.DeviceArray.from_host(ctx, y0)
        F = gg.Frontier.from_dense(g, front) if front is not None else gg.Frontier.bits(g)
        out, path = gg.edge_map(g, F, k, a, b)
        assert path == gg.PATH_COO
        got.append((out.dense(), b.to_host()))
        out.close()
        F.close()
    assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    act = np.ones(len(ids), bool) if front is None else front[src].astype(bool)
    od = np.diff(off.astype(np.int64)).astype(np.float64)
    xs = x.astype(np.float64)[src]
    term = xs / od[src] if kind == "pagerank" else xs * (w.astype(np.float64) if kind == "spmv" else 1.0)
    want = y0.astype(np.float64) + np.bincount(ids[act], term[act], minlength=n)
    flags = np.zeros(n, np.uint8)
    flags[ids[act]] = 1
    assert np.array_equal(got[0][0], flags)
    assert _rel(got[0][1], want) < tol
    g.close()


@pytest.mark.parametrize("scale,P", [(14, 64), (18, 384)])
def test_coo_pagerank_loop_vs_oracle(ctx, scale, P):
    """PageRank.C with every dense step on the COO traversal, through the
    generic edgeMap/vertexMap loop and the fused loop, against the oracle's
    COO run (its PART96 sequential adds in Hilbert order); bitwise equal run
    to run."""
    off, ids = oracle.rmat_csr(scale, 16, 1)
    n = len(off) - 1
    og = oracle.Graph(off, ids, P=P)
    want, it, tr = og.pagerank(10)
    g = gg.Graph.from_csr(ctx, off, ids).partition(P)
    g.set_option(gg.OPT_DENSE_COO, 1)
    runs = []
    for _ in range(2):
        r, info = gg.pagerank(g, np.float64)
        assert info.iters == it
        runs.append(r.to_host())
    assert np.array_equal(runs[0], runs[1])
    assert R.rel_l1(runs[0], want) < 1e-12
    r32, _ = gg.pagerank(g, np.float32)
    assert R.rel_l1(r32.to_host(), want) < 1e-5
    pc = gg.DeviceArray(ctx, n, np.float64).fill(1.0 / n)
    pn = gg.DeviceArray(ctx, n, np.float64).fill(0.0)
    F = gg.Frontier.bits(g)
    paths = []
    for _ in range(10):
        out, path = gg.edge_map(g, F, gg.OP_PAGERANK, pc, pn, threshold=len(ids) // 20)
        paths.append(path)
        gg.vertex_map(g, F, gg.VOP_PAGERANK, pc, pn, 0.85, (1 - 0.85) * (1 / n))
        gg.vertex_map(g, None, gg.VOP_ABSDIFF, pc, pn)
        gg.vertex_map(g, F, gg.VOP_RESET, pc)
        pc, pn = pn, pc
        F.close()
        F = out
    F.close()
    assert paths == [t.path for t in tr] and gg.PATH_COO in paths
    assert R.rel_l1(pc.to_host(), want) < 1e-12
    g.close()


def test_coo_spmv_vs_reference_golden(ctx):
    """SPMV.C (integer weights, x = 1/n or the fixture's x) on the COO
    traversal against the reference engine's own result."""
    for path in R.golden_cases():
        z = R.load_golden(path)
        if "spmv_edge" not in z:
            continue
        w = z["spmv_w"].astype(np.float32)
        g = gg.Graph.from_csr(ctx, z["off"], z["ids"], w=w).partition(int(z["P"]))
        g.set_option(gg.OPT_DENSE_COO, 1)
        n = g.n
        x = z["spmv_x"] if "spmv_x" in z else np.full(n, 1.0 / n)
        xd = gg.DeviceArray.from_host(ctx, x.astype(np.float64))
        y = gg.DeviceArray(ctx, n, np.float64).fill(0.0)
        F = gg.Frontier.bits(g)
        out, pth = gg.edge_map(g, F, gg.OP_SPMV, xd, y)
        assert pth == gg.PATH_COO
        assert R.rel_l1(y.to_host(), z["spmv_edge"]) < 1e-12, path
        out.close()
        F.close()
        g.close()


def test_coo_option_values(ctx):
    g, _, _, _ = _graph(ctx, 8, 1)
    assert g.get_option(gg.OPT_DENSE_COO) == 0
    g.set_option(gg.OPT_DENSE_COO, 1)
    assert g.get_option(gg.OPT_DENSE_COO) == 1
    with pytest.raises(gg.GGError):
        g.set_option(gg.OPT_DENSE_COO, 2)
    g.close()


@pytest.mark.parametrize("sym", [True, False])
def test_coo_cc_bfs_pathsum(ctx, sym):
    g, off, ids, _ = _graph(ctx, 12, 4, sym=sym)
    n = g.n
    rng = np.random.default_rng(9)
    front = (rng.random(n) < 0.8).astype(np.uint8)
    # Components.C one round: labels and the changed set exact
    lab0 = rng.permutation(n).astype(np.uint32)
    (f0, a0, _, _), (f1, a1, _, _) = _both(
        g, front, gg.OP_CC, lambda: (gg.DeviceArray.from_host(ctx, lab0), gg.DeviceArray.from_host(ctx, lab0)))
    assert np.array_equal(a0, a1) and np.array_equal(f0, f1)
    # BFS.C one level: the same newly visited set, every new parent a frontier
    # in-neighbour
    par0 = np.where(rng.random(n) < 0.3, np.arange(n), -1).astype(np.int32)
    (f0, p0, _, _), (f1, p1, _, _) = _both(g, front, gg.OP_BFS, lambda: (gg.DeviceArray.from_host(ctx, par0), None))
    assert np.array_equal(f0, f1) and np.array_equal(p0 >= 0, p1 >= 0)
    new = np.flatnonzero(f1)
    src = np.repeat(np.arange(n), np.diff(off))
    edges = set(zip(src.tolist(), ids.tolist()))
    for d in new[:2000]:
        assert front[p1[d]] and (int(p1[d]), int(d)) in edges
    # BC.C forward step: path counts (integers in FP64) exact
    vis = (rng.random(n) < 0.3).astype(np.uint8) | front
    x0 = np.where(front.astype(bool), rng.integers(1, 9, n), 0).astype(np.float64)
    (f0, x0a, _, _), (f1, x1a, _, _) = _both(
        g, front, gg.OP_PATHSUM, lambda: (gg.DeviceArray.from_host(ctx, x0), gg.DeviceArray.from_host(ctx, vis)))
    assert np.array_equal(f0, f1) and np.array_equal(x0a, x1a)
    g.close()


def test_coo_sssp_whole(ctx):
    """BellmanFord.C with the COO traversal: one round may relax in another
    order than the pull (the reference's COO loop reads live distances), so
    the check is on the converged distances, exact."""
    off, ids = oracle.rmat_csr(12, 16, 6)
    wl = (1 + (np.arange(len(ids)) * 2654435761 % 19)).astype(np.float32)
    g = gg.Graph.from_csr(ctx, off, ids, w=wl).partition(16)
    want, _ = oracle.sssp(off, ids, wl, 0)
    for opt in (0, 1):
        g.set_option(gg.OPT_DENSE_COO, opt)
        d, _ = gg.sssp(g, 0)
        assert np.array_equal(d.to_host().astype(np.int64), want), opt
    g.close()


@pytest.mark.parametrize("sym", [True, False])
def test_coo_apps_whole(ctx, sym):
    """BFS, Components, BC and Bellman-Ford on a partitioned graph with the
    COO traversal on its COO levels: the oracle's answers."""
    off, ids = oracle.rmat_csr(14, 16, 2, symmetric=sym)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=sym).partition(64)
    g.set_option(gg.OPT_DENSE_COO, 1)
    og = oracle.Graph(off, ids, symmetric=sym, P=64)
    _, lv, tr = og.bfs(100)
    p, l, info = gg.bfs(g, 100)
    R.check_parents(off, ids, 100, p.to_host())
    assert np.array_equal(l.to_host(), R.bfs_levels(off, ids, 100)) and info.iters == lv
    want_cc, _, _ = og.cc()
    lab, cinfo = gg.components(g)
    assert np.array_equal(lab.to_host(), want_cc)
    want_bc, wl, _, rounds = oracle.bc(off, ids, 100)
    sc, blv, binfo = gg.bc(g, 100, np.float64)
    reached = wl >= 0
    got = sc.to_host()
    assert np.array_equal(blv.to_host(), wl) and binfo.iters == rounds
    assert _rel(got[reached], want_bc[reached]) < 1e-12
    paths = info.path_list() + cinfo.path_list() + binfo.path_list()
    assert "coo" in paths  # the COO traversal ran
    g.close()
