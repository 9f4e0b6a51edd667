"""GPU parity of the batched forms: the arena graph, personalised PageRank
and its adjoint over many packed graphs (captured fixed-iteration loop),
32-source BFS, and the pull autotuner.

Bars: arena layout bit-exact; per-graph PPR/adjoint slices <= 1e-12 rel L1
(FP64) and <= 1e-5 (FP32) against the FP64 oracle run graph by graph;
captured replays bitwise equal to the eager loop; MS-BFS levels bit-exact
against textbook BFS per source.
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle
import refcheck as R

pytestmark = pytest.mark.gpu

ALPHA = 0.85


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _parts(ctx, specs, weighted=True):
    gs = [gg.Graph.rmat(ctx, s, ef, seed, weighted=weighted) for (s, ef, seed) in specs]
    return gs


SPECS = [(8, 8, 51), (10, 16, 52), (6, 32, 53), (12, 8, 54), (9, 16, 55), (11, 32, 56), (7, 8, 57)]


def test_arena_layout(ctx):
    parts = _parts(ctx, SPECS)
    a, vb, eb = gg.Graph.arena(ctx, parts)
    assert a.n == sum(p.n for p in parts) and a.m == sum(p.m for p in parts)
    off, ids, w = a.csr()
    for i, p in enumerate(parts):
        po, pi, pw = p.csr()
        assert vb[i] == sum(q.n for q in parts[:i]) and eb[i] == sum(q.m for q in parts[:i])
        assert np.array_equal(off[vb[i]:vb[i + 1] + 1] - eb[i], po)
        assert np.array_equal(ids[eb[i]:eb[i + 1]] - vb[i], pi)
        assert np.array_equal(w[eb[i]:eb[i + 1]], pw)
    # the in-edge twin of the union is the union of the twins
    toff, tids, _ = a.csc()
    for i, p in enumerate(parts):
        qo, qi, _ = p.csc()
        assert np.array_equal(toff[vb[i]:vb[i + 1] + 1] - eb[i], qo)
        assert np.array_equal(tids[eb[i]:eb[i + 1]] - vb[i], qi)
    a.close()
    for p in parts:
        p.close()


def _seeds(vb, seed):
    """1 to 8 seeds per graph; v uniform over each graph's seeds."""
    rng = np.random.default_rng(seed)
    n = int(vb[-1])
    v = np.zeros(n)
    for i in range(len(vb) - 1):
        ni = int(vb[i + 1] - vb[i])
        k = int(rng.integers(1, 9))
        s = rng.choice(ni, min(k, ni), replace=False)
        v[vb[i] + s] += 1.0 / len(s)
    return v


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_batched_ppr_and_adjoint_vs_oracle(ctx, dt, tol):
    print(f"batched ppr {np.dtype(dt).name}", flush=True)
    parts = _parts(ctx, SPECS, weighted=False)
    a, vb, eb = gg.Graph.arena(ctx, parts)
    ones = gg.DeviceArray(ctx, a.m, np.float32).fill(1.0)
    a.set_weights_dev(ones)
    v = _seeds(vb, 5)
    vd = gg.DeviceArray.from_host(ctx, v.astype(dt))
    r, info = gg.wpagerank(a, dt, ALPHA, -1.0, 30, v=vd)  # fixed 30 iterations, captured
    assert info.iters == 30
    rh = r.to_host()
    # loss: sum of r^2 over 8 targets per graph
    rng = np.random.default_rng(6)
    u = np.zeros(a.n)
    for i in range(len(parts)):
        t = vb[i] + rng.choice(int(vb[i + 1] - vb[i]), 8, replace=False)
        u[t] = 2 * rh[t]
    ud = gg.DeviceArray.from_host(ctx, u.astype(dt))
    gw, gv, lam, ainfo = gg.wpagerank_adjoint(a, r, ud, ALPHA, -1.0, 30, want_lam=True)
    assert ainfo.iters == 30
    gwh, gvh = gw.to_host()[: a.m], gv.to_host()
    for i, p in enumerate(parts):
        po, pi, _ = p.csr()
        w1 = np.ones(len(pi), np.float32)
        sl, el = slice(vb[i], vb[i + 1]), slice(eb[i], eb[i + 1])
        want, it, _ = oracle.wpagerank(po, pi, w1, ALPHA, 0.0, 30, v=v[sl])
        assert it == 30
        assert _rel(rh[sl], want) < tol, (i, _rel(rh[sl], want))
        _, gw_w, gv_w, ait = oracle.wpagerank_adjoint(po, pi, w1, rh[sl].astype(np.float64), u[sl], ALPHA, 0.0, 30)
        assert ait == 30
        assert _rel(gvh[sl], gv_w) < tol
        if len(pi):
            assert _rel(gwh[el], gw_w) < tol * 10  # differences of two nearby values
    # the captured replays equal the eager loop bit for bit, replay after replay
    r_e, _ = gg.wpagerank(a, dt, ALPHA, 0.0, 30, v=vd)
    gw_e, gv_e, _, _ = gg.wpagerank_adjoint(a, r_e, ud, ALPHA, 0.0, 30)
    for _ in range(3):
        ctx.sync()
        r2, _ = gg.wpagerank(a, dt, ALPHA, -1.0, 30, v=vd, out=r)
        assert np.array_equal(r2.to_host(), r_e.to_host())
        _ = gg.DeviceArray(ctx, 1024, np.float64).fill(3.0)  # eager work between replays
        gw2, gv2, _, _ = gg.wpagerank_adjoint(a, r, ud, ALPHA, -1.0, 30, gw=gw, gv=gv, want_lam=False)
        assert np.array_equal(gw2.to_host()[: a.m], gw_e.to_host()[: a.m])
        assert np.array_equal(gv2.to_host(), gv_e.to_host())
    a.close()
    for p in parts:
        p.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_msbfs_levels_exact(ctx, plan):
    """plan 1: the levels run on the hot-sources-first renumbering (sources
    mapped in, per-source levels mapped back)."""
    g = gg.Graph.rmat(ctx, 14, 16, 2, symmetric=True).reorder(plan)
    off, ids, _ = g.csr()
    n = g.n
    deg = np.diff(off.astype(np.int64))
    rng = np.random.default_rng(7)
    src = rng.choice(np.flatnonzero(deg > 0), 29, replace=False).tolist()
    src += [src[0], int(np.flatnonzero(deg == 0)[0]), 100]  # a repeat, an isolated vertex
    lv, info = gg.msbfs(g, src)
    L = lv.to_host().reshape(32, n)
    for k, s in enumerate(src):
        assert np.array_equal(L[k], R.bfs_levels(off, ids, s)), k
    assert info.iters == int(L.max()) + 1  # the last level finds nothing new
    # level 1 pushes from the 32 sources (a symmetric graph)
    names = info.path_list()
    assert len(names) == info.iters and names[0] == "sparse" and set(names) <= {"sparse", "csc"}, names
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_msbfs_directed_and_limits(ctx, plan):
    g = gg.Graph.rmat(ctx, 12, 8, 3).reorder(plan)
    off, ids, _ = g.csr()
    src = [0, 5, 17]
    lv, info = gg.msbfs(g, src)
    assert set(info.path_list()) == {"csc"}  # a directed graph always pulls
    L = lv.to_host().reshape(3, g.n)
    for k, s in enumerate(src):
        assert np.array_equal(L[k], R.bfs_levels(off, ids, s))
    with pytest.raises(gg.GGError):
        gg.msbfs(g, list(range(33)))
    with pytest.raises(gg.GGError):
        gg.msbfs(g, [g.n])
    g.close()


def test_msbfs_scale22_against_single_source(ctx):
    """The multi-source BFS shape (symmetric R-MAT scale 22, ef 16, 32
    sources): three sources against the textbook BFS, all 32 against the
    single-source engine's levels on a sample of sources, and the level
    structure of every source (neighbours differ by at most one level)."""
    print("msbfs scale=22", flush=True)
    g = gg.Graph.rmat(ctx, 22, 16, 2, symmetric=True).reorder(1)
    off, ids, _ = g.csr()
    n = g.n
    deg = np.diff(off.astype(np.int64))
    src = np.random.default_rng(22).choice(np.flatnonzero(deg > 0), 32, replace=False).tolist()
    lv, info = gg.msbfs(g, src)
    names = info.path_list()
    assert "sparse" in names and "csc" in names, names  # both branches at this size
    L = lv.to_host().reshape(32, n)
    for k in (0, 13, 31):
        assert np.array_equal(L[k], R.bfs_levels(off, ids, src[k])), k
    g.partition(384)
    for k in (1, 7, 20):
        _, l1, _ = gg.bfs(g, src[k])
        assert np.array_equal(L[k], l1.to_host()), k
    srcv = np.repeat(np.arange(n, dtype=np.int64), deg)
    for k in range(0, 32, 8):
        a, b = L[k][srcv], L[k][ids]
        both = (a >= 0) & (b >= 0)
        assert np.all(np.abs(a[both] - b[both]) <= 1) and np.all((a >= 0) == (b >= 0))
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_msbfs_many_levels_and_depth_changes(ctx, plan):
    """A symmetric path with shuffled ids: levels are enqueued in batches
    (the first sized by the previous call's depth), so calls whose depth
    needs several batches, and calls shallower than the previous one (the
    levels past the end enqueued and skipped), must give the same levels and
    path counts as the textbook BFS."""
    n = 230
    order = np.random.default_rng(11).permutation(n)
    off, ids = oracle.csr_from_edges(n, order[:-1].astype(np.uint32), order[1:].astype(np.uint32),
                                     symmetric=True)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).partition(8).reorder(plan)
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    # path positions of the sources: an end (depth n), the middle, near the middle
    for where in ([0], [0, 115], [115], [100, 130, 115], [229, 0]):
        src = [int(order[p]) for p in where]
        print(f"msbfs path n={n} plan={plan} sources at {where}", flush=True)
        depth = max(max(p, n - 1 - p) for p in where)
        lv, info = gg.msbfs(g, src)
        L = lv.to_host().reshape(len(src), n)
        for k, s in enumerate(src):
            assert np.array_equal(L[k], R.bfs_levels(off, ids, s)), (where, k)
        assert info.iters == depth + 1
        # frontiers of one or two vertices push (the first 64 levels are logged)
        assert info.path_list() == ["sparse"] * min(info.iters, 64)
        pa, lv2, binfo = gg.msbc_forward(g, src, np.float64, levels=True)
        assert np.array_equal(lv2.to_host().reshape(len(src), n), L)
        assert binfo.iters == depth + 1
        # one shortest path to every vertex of a path
        assert np.array_equal(pa.to_host().reshape(len(src), n), np.ones((len(src), n)))
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_msbfs_fewer_vertices_than_sources(ctx, plan):
    """32 sources on a 6-vertex symmetric graph (every source repeated, an
    isolated vertex among them): the levels and path counts of each."""
    off, ids = oracle.csr_from_edges(6, np.array([0, 1, 2, 0], np.uint32), np.array([1, 2, 3, 2], np.uint32),
                                     symmetric=True)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).reorder(plan)
    src = [k % 6 for k in range(32)]
    print(f"msbfs n=6 plan={plan}", flush=True)
    lv, _ = gg.msbfs(g, src)
    L = lv.to_host().reshape(32, 6)
    pa, lv2, _ = gg.msbc_forward(g, src, np.float64, levels=True)
    P = pa.to_host().reshape(32, 6)
    for k, s in enumerate(src):
        _, level, sigma, _ = oracle.bc(off, ids, s)
        assert np.array_equal(L[k], R.bfs_levels(off, ids, s)), k
        assert np.array_equal(L[k], level), k
        assert np.array_equal(P[k], np.where(level >= 0, sigma, 0.0)), k
    assert np.array_equal(lv2.to_host().reshape(32, 6), L)
    g.close()


@pytest.mark.parametrize("scale,ef", [(13, 16), (15, 8)])
def test_msbfs_segmented_pull(ctx, scale, ef):
    """The source-segmented layout forced on (it is taken from 2^24 edges by
    default): dense levels where most rows still gather pull through it.
    Levels and FP64 path counts of all 32 sources against the oracle, and the
    same results with the layout off."""
    g = gg.Graph.rmat(ctx, scale, ef, 9, symmetric=True).reorder(1)
    g.set_option(3, 2)  # GG_OPT_XCD_SEGMENTS: always
    off, ids, _ = g.csr()
    n = g.n
    deg = np.diff(off.astype(np.int64))
    src = np.random.default_rng(scale).choice(np.flatnonzero(deg > 0), 32, replace=False).tolist()
    print(f"msbfs segmented scale={scale} ef={ef}", flush=True)
    lv, _ = gg.msbfs(g, src)
    pa, lv2, _ = gg.msbc_forward(g, src, np.float64, levels=True)
    L, P = lv.to_host().reshape(32, n), pa.to_host().reshape(32, n)
    for k, s in enumerate(src):
        _, level, sigma, _ = oracle.bc(off, ids, s)
        assert np.array_equal(L[k], level), k
        assert np.array_equal(P[k], np.where(level >= 0, sigma, 0.0)), k
    assert np.array_equal(lv2.to_host().reshape(32, n), L)
    g.set_option(3, 0)
    lv3, _ = gg.msbfs(g, src)
    assert np.array_equal(lv3.to_host().reshape(32, n), L)
    g.close()


def test_autotune_picks_fastest_and_keeps_results(ctx):
    g = gg.Graph.rmat(ctx, 16, 16, 4, weighted=True)
    t = gg.autotune_pull(g, np.float64, trials=3)
    assert t["candidates"] == 7
    ms = t["cand_ms"]
    assert t["ms"] == min(ms)
    assert g.get_option(gg.OPT_PULL_SCHEDULE) == t["schedule"]
    assert g.get_option(gg.OPT_ROW_EDGES_PER_LANE) == t["edges_per_lane"]
    off, ids, w = g.csr()
    want, it, _ = oracle.wpagerank(off, ids, w, ALPHA, 1e-7, 100)
    r, info = gg.wpagerank(g, np.float64, ALPHA, 1e-7, 100)
    assert info.iters == it and _rel(r.to_host(), want) < 1e-12
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale,sym", [(12, True), (13, False)])
def test_msbc_forward_paths_exact(ctx, scale, sym, plan):
    """BC.C's forward phase for 32 sources at once (a repeat and an isolated
    vertex among them): every source's levels and NumPaths equal the
    oracle's single-source BC (path counts are integers, exact in FP64;
    FP32 within its rounding)."""
    print(f"msbc scale={scale} sym={sym} plan={plan}", flush=True)
    g = gg.Graph.rmat(ctx, scale, 16, 2, symmetric=sym).reorder(plan)
    off, ids, _ = g.csr()
    n = g.n
    deg = np.diff(off.astype(np.int64))
    rng = np.random.default_rng(11)
    src = rng.choice(np.flatnonzero(deg > 0), 30, replace=False).tolist()
    src += [src[0], int(np.flatnonzero(deg == 0)[0])]
    pa, lv, info = gg.msbc_forward(g, src, np.float64, levels=True)
    P = pa.to_host().reshape(32, n)
    L = lv.to_host().reshape(32, n)
    for k, s in enumerate(src):
        _, level, sigma, _ = oracle.bc(off, ids, s)
        assert np.array_equal(L[k], level), k
        assert np.array_equal(P[k], np.where(level >= 0, sigma, 0.0)), k
    p32, _, _ = gg.msbc_forward(g, src, np.float32)
    P32 = p32.to_host().reshape(32, n).astype(np.float64)
    assert np.all(np.abs(P32 - P) <= 1e-6 * np.maximum(np.abs(P), 1.0))
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_msbc_forward_hubs(ctx, plan):
    """Vertices with in-lists longer than one 1024-edge chunk (their path
    counts are summed in chunks, then in chunk order): a symmetric R-MAT
    graph with two added hubs of 3000 and 5000 neighbours."""
    off, ids = oracle.rmat_csr(13, 8, 5, symmetric=True)
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    rng = np.random.default_rng(3)
    extra = []
    for hub, k in ((7, 3000), (1000, 5000)):
        nb = rng.choice(n, k, replace=False)
        nb = nb[nb != hub]
        extra += [(hub, int(u)) for u in nb] + [(int(u), hub) for u in nb]
    es, ed = zip(*extra)
    off2, ids2 = oracle.csr_from_edges(n, np.concatenate([src, es]), np.concatenate([ids.astype(np.int64), ed]))
    g = gg.Graph.from_csr(ctx, off2, ids2, symmetric=True).reorder(plan)
    deg = np.diff(off2.astype(np.int64))
    sources = [int(v) for v in rng.choice(np.flatnonzero(deg > 0), 20, replace=False)] + [7, 1000]
    pa, lv, _ = gg.msbc_forward(g, sources, np.float64, levels=True)
    P = pa.to_host().reshape(len(sources), n)
    L = lv.to_host().reshape(len(sources), n)
    for k, s in enumerate(sources):
        _, level, sigma, _ = oracle.bc(off2, ids2, s)
        assert np.array_equal(L[k], level), k
        assert np.array_equal(P[k], np.where(level >= 0, sigma, 0.0)), k
    g.close()


def test_msbc_forward_limits(ctx):
    g = gg.Graph.rmat(ctx, 10, 8, 3)
    with pytest.raises(gg.GGError):
        gg.msbc_forward(g, list(range(33)))
    with pytest.raises(gg.GGError):
        gg.msbc_forward(g, [g.n])
    with pytest.raises(ValueError):
        gg.msbc_forward(g, [0], np.int32)
    g.close()
