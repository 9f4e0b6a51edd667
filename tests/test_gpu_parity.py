"""GPU parity: the HIP path through the C ABI against the pinned oracle.

Integer and index results (graph layout, partitions, COO order, BFS levels,
CC labels, frontier statistics) must match bit-exactly; floating-point
results within the stated tolerances: rel L1 <= 1e-12 in FP64 and <= 1e-5
in FP32 against the FP64 oracle.
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle
import refcheck as R

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _dev_graph(ctx, off, ids, symmetric=False, w=None):
    return gg.Graph.from_csr(ctx, off, ids, w=w, symmetric=symmetric)


# ---------------------------------------------------------------- layout --

@pytest.mark.parametrize("scale,ef,seed,sym", [(8, 16, 1, False), (12, 16, 2, True),
                                               (14, 8, 3, False), (16, 16, 1, False)])
def test_rmat_generator_matches_oracle(ctx, scale, ef, seed, sym):
    print(f"rmat scale={scale} ef={ef} seed={seed} sym={sym}", flush=True)
    g = gg.Graph.rmat(ctx, scale, ef, seed, symmetric=sym)
    off, ids = oracle.rmat_csr(scale, ef, seed, symmetric=sym)
    goff, gids, _ = g.csr()
    assert g.m == len(ids)
    assert np.array_equal(goff, off) and np.array_equal(gids, ids)
    toff, tids, _ = oracle.transpose(off, ids)
    coff, cids, _ = g.csc()
    assert np.array_equal(coff, toff) and np.array_equal(cids, tids)
    assert np.array_equal(g.degrees(False), np.diff(off))
    assert np.array_equal(g.degrees(True), np.diff(toff))
    g.close()


def test_rmat_weights_match_oracle(ctx):
    g = gg.Graph.rmat(ctx, 10, 16, 3, weighted=True)
    off, ids, w = g.csr()
    assert np.array_equal(w, oracle.weights(3, len(ids)))
    toff, tids, perm = oracle.transpose(off, ids)
    _, _, tw = g.csc()
    assert np.array_equal(tw, w[perm])
    g.close()


@pytest.mark.parametrize("path", R.golden_cases(), ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("by_vertex", [False, True])
def test_partitions_match_golden(ctx, path, by_vertex):
    z = R.load_golden(path)
    tag = "vertex" if by_vertex else "edge"
    P = int(z["P"])
    g = _dev_graph(ctx, z["off"], z["ids"], bool(z["symmetric"])).partition(P, by_vertex)
    s0, _ = g.starts(0)
    s1, _ = g.starts(1)
    assert np.array_equal(s0, z[f"coo_starts_{tag}"])
    assert np.array_equal(s1, z[f"csc_starts_{tag}"])
    if not by_vertex:
        og = oracle.Graph(z["off"], z["ids"], symmetric=bool(z["symmetric"]), P=P)
        src, dst, po = g.coo()
        for p in range(P):
            s, d = og.coo(p)
            assert np.array_equal(src[po[p]:po[p + 1]], s) and np.array_equal(dst[po[p]:po[p + 1]], d)
    g.close()


@pytest.mark.parametrize("scale,P", [(14, 64), (16, 384), (20, 384)])
def test_partitions_rmat_match_oracle(ctx, scale, P):
    print(f"partition scale={scale} P={P}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 1)
    og = oracle.Graph(off, ids, P=P)
    g = _dev_graph(ctx, off, ids).partition(P)
    s0, under0 = g.starts(0)
    assert np.array_equal(s0, og.starts(0))
    assert np.array_equal(g.starts(1)[0], og.starts(1))
    assert under0 == bool(oracle.lib().ggo_graph_error() & 1)
    src, dst, po = g.coo()
    for p in range(0, P, max(1, P // 16)):
        s, d = og.coo(p)
        assert np.array_equal(src[po[p]:po[p + 1]], s) and np.array_equal(dst[po[p]:po[p + 1]], d)
    g.close()


def test_partition_overrun_is_an_error(ctx):
    # same degree sequence as the oracle's overrun case: in-degrees 3,1,2,3,3
    src = [1, 2, 3, 0, 0, 1, 0, 1, 2, 0, 1, 2]
    dst = [0, 0, 0, 1, 2, 2, 3, 3, 3, 4, 4, 4]
    off, ids = oracle.csr_from_edges(5, src, dst)
    g = _dev_graph(ctx, off, ids)
    with pytest.raises(gg.GGError) as e:
        g.partition(3)
    assert e.value.code == -5
    g.close()


# -------------------------------------------------------------- PageRank --

def _pr_cases():
    return [p for p in R.golden_cases() if "pagerank_edge" in np.load(p).files]


@pytest.mark.parametrize("path", _pr_cases(), ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("by_vertex", [False, True])
def test_pagerank_matches_reference_golden(ctx, path, by_vertex):
    z = R.load_golden(path)
    tag = "vertex" if by_vertex else "edge"
    g = _dev_graph(ctx, z["off"], z["ids"]).partition(int(z["P"]), by_vertex)
    r, info = gg.pagerank(g, np.float64)
    assert info.iters == int(z[f"pagerank_{tag}_iters"])
    assert R.rel_l1(r.to_host(), z[f"pagerank_{tag}"]) < TOL64
    g.close()


@pytest.mark.parametrize("scale,P", [(14, 64), (18, 384), (20, 384)])
def test_pagerank_fp64_fp32_vs_oracle(ctx, scale, P):
    print(f"pagerank scale={scale} P={P}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 1)
    og = oracle.Graph(off, ids, P=P)
    want, it, tr = og.pagerank(10)
    g = _dev_graph(ctx, off, ids).partition(P)
    r64, info = gg.pagerank(g, np.float64)
    assert info.iters == it
    assert info.path_list() == [oracle_path(t.path) for t in tr]
    assert R.rel_l1(r64.to_host(), want) < TOL64
    r32, _ = gg.pagerank(g, np.float32)
    assert R.rel_l1(r32.to_host(), want) < TOL32
    # fixed summation order: a second run is bitwise identical
    again, _ = gg.pagerank(g, np.float64)
    assert np.array_equal(again.to_host().view(np.uint64), r64.to_host().view(np.uint64))
    g.close()


def _check_row_blocks(g):
    """The row-block table of the renumbered plan: every in-edge in exactly
    one block, blocks of complete rows within the staging limits, long rows
    chunked, every row finished once."""
    v = g.device_view(1)
    assert v.row_work and v.row_blocks > 0
    tab = np.empty(8 * v.row_blocks, np.uint32)
    gg.check(gg.lib().gg_memcpy_d2h(g.ctx.h, tab.ctypes.data, v.row_work, tab.nbytes))
    tab = tab.reshape(-1, 8).astype(np.int64)
    toff = np.empty(g.n + 1, np.uint32)
    gg.check(gg.lib().gg_memcpy_d2h(g.ctx.h, toff.ctypes.data, v.toff, toff.nbytes))
    toff = toff.astype(np.int64)
    r0, r1, e0, e1, kind = tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], tab[:, 4]
    assert np.all((e1 - e0) <= 2048) and np.all(e1 >= e0)
    # blocks tile the edge range in order
    assert e0[0] == 0 and e1[-1] == g.m and np.array_equal(e0[1:], e1[:-1])
    short = kind != 0xFF
    assert np.all(kind[short] <= 6) and np.all((r1 - r0)[short] >= 1) and np.all((r1 - r0)[short] <= 1024)
    assert np.array_equal(toff[r0[short]], e0[short]) and np.array_equal(toff[r1[short]], e1[short])
    rows = np.zeros(g.n, np.int64)
    np.add.at(rows, np.concatenate([np.arange(a, b) for a, b in zip(r0[short], r1[short])]), 1)
    lrows = np.unique(r0[~short])
    assert len(lrows) == v.row_long and np.all(np.diff(toff)[lrows] > 2048)
    rows[lrows] += 1
    assert np.all(rows == 1)


@pytest.mark.parametrize("scale,P", [(12, 64), (20, 384)])
@pytest.mark.parametrize("schedule", [1, 0])
def test_pagerank_reordered_plan(ctx, scale, P, schedule):
    """The hot-sources-first renumbering changes only summation order; both
    pull schedules (row blocks, merge-path tiles) run it."""
    print(f"pagerank reordered scale={scale} schedule={schedule}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 1)
    og = oracle.Graph(off, ids, P=P)
    want, it, _ = og.pagerank(10)
    g = _dev_graph(ctx, off, ids).partition(P).reorder(1)
    g.set_option(gg.OPT_PULL_SCHEDULE, schedule)
    if schedule == 1:
        _check_row_blocks(g)
    r64, info = gg.pagerank(g, np.float64)
    assert info.iters == it
    assert R.rel_l1(r64.to_host(), want) < TOL64
    r32, _ = gg.pagerank(g, np.float32)
    assert R.rel_l1(r32.to_host(), want) < TOL32
    again, _ = gg.pagerank(g, np.float64)
    assert np.array_equal(again.to_host().view(np.uint64), r64.to_host().view(np.uint64))
    g.reorder(0)
    plain, _ = gg.pagerank(g, np.float64)
    assert R.rel_l1(plain.to_host(), want) < TOL64
    g.close()


def oracle_path(p):
    return {oracle.PATH_SPARSE: "sparse", oracle.PATH_CSC: "csc", oracle.PATH_COO: "coo"}[p]


def test_pagerank_break_on_convergence(ctx):
    # a 2-cycle converges: L1 drops below 1e-7 before 100 iterations
    off, ids = oracle.csr_from_edges(2, [0, 1], [1, 0])
    og = oracle.Graph(off, ids, P=1)
    want, it, _ = og.pagerank(100)
    g = _dev_graph(ctx, off, ids).partition(1)
    r, info = gg.pagerank(g, np.float64, max_iter=100)
    assert info.iters == it < 100
    assert R.rel_l1(r.to_host(), want) < TOL64
    g.close()


def test_pagerank_generic_edgemap_loop(ctx):
    """PageRank.C written against the edgeMap/vertexMap entry points."""
    off, ids = oracle.rmat_csr(12, 16, 1)
    n = len(off) - 1
    og = oracle.Graph(off, ids, P=64)
    want, it, tr = og.pagerank(10)
    g = _dev_graph(ctx, off, ids).partition(64)
    pc = gg.DeviceArray(ctx, n, np.float64).fill(1.0 / n)
    pn = gg.DeviceArray(ctx, n, np.float64).fill(0.0)
    F = gg.Frontier.bits(g)
    paths = []
    for count in range(10):
        out, path = gg.edge_map(g, F, gg.OP_PAGERANK, pc, pn, threshold=len(ids) // 20)
        paths.append(gg.PATH_NAMES[path])
        gg.vertex_map(g, F, gg.VOP_PAGERANK, pc, pn, 0.85, (1 - 0.85) * (1 / n))
        gg.vertex_map(g, None, gg.VOP_ABSDIFF, pc, pn)
        assert gg.reduce_sum(ctx, pc) >= 1e-7
        gg.vertex_map(g, F, gg.VOP_RESET, pc)
        pc, pn = pn, pc
        F.close()
        F = out
    assert paths == [oracle_path(t.path) for t in tr]
    assert R.rel_l1(pc.to_host(), want) < TOL64
    g.close()


# ------------------------------------------------------------ BFS and CC --

@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale,sym,root", [(12, True, 100), (16, True, 100), (16, False, 3)])
def test_bfs_levels_exact(ctx, scale, sym, root, plan):
    """plan 1: the levels run on the hot-sources-first renumbering and
    parents and levels are mapped back; the branch taken at every level is
    the oracle's either way (the frontier statistics do not depend on the
    numbering)."""
    print(f"bfs scale={scale} sym={sym} plan={plan}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 2, symmetric=sym)
    og = oracle.Graph(off, ids, symmetric=sym, P=64)
    opar, lv, tr = og.bfs(root)
    g = _dev_graph(ctx, off, ids, sym).partition(64).reorder(plan)
    p, l, info = gg.bfs(g, root)
    par = p.to_host()
    lvl = R.check_parents(off, ids, root, par)
    assert np.array_equal(l.to_host(), lvl)
    assert info.iters == lv
    assert info.path_list() == [oracle_path(t.path) for t in tr]
    g.close()


def test_bfs_plan_without_levels(ctx):
    """The plan's map-back when the caller asks for parents only."""
    off, ids = oracle.rmat_csr(13, 8, 7, symmetric=True)
    g = _dev_graph(ctx, off, ids, True).reorder(1)
    p, l, _ = gg.bfs(g, 5, levels=False)
    assert l is None
    R.check_parents(off, ids, 5, p.to_host())
    g.close()


@pytest.mark.parametrize("path", [p for p in R.golden_cases() if "bfs_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
def test_bfs_matches_reference_golden(ctx, path):
    z = R.load_golden(path)
    root = int(z["root"])
    g = _dev_graph(ctx, z["off"], z["ids"], bool(z["symmetric"])).partition(int(z["P"]))
    for plan in (0, 1):
        g.reorder(plan)
        p, l, info = gg.bfs(g, root)
        lvl = R.check_parents(z["off"], z["ids"], root, p.to_host())
        assert np.array_equal(l.to_host(), lvl)
        R.check_parents(z["off"], z["ids"], root, z["bfs_edge"].astype(np.int64), lvl)
        assert info.iters == int(z["bfs_edge_iters"])
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale,sym", [(12, True), (16, True), (14, False)])
def test_cc_labels_exact(ctx, scale, sym, plan):
    """plan 1: the loop runs on the hot-sources-first renumbering with the
    caller's ids as label values; the fixed point is the same."""
    print(f"cc scale={scale} sym={sym} plan={plan}", flush=True)
    off, ids = oracle.rmat_csr(scale, 16, 2, symmetric=sym)
    og = oracle.Graph(off, ids, symmetric=sym, P=64)
    want, rounds, _ = og.cc()
    g = _dev_graph(ctx, off, ids, sym).partition(64).reorder(plan)
    lab, info = gg.components(g)
    assert np.array_equal(lab.to_host(), want)
    if sym:
        assert np.array_equal(lab.to_host(), R.cc_min_labels(off, ids))
    g.close()


@pytest.mark.parametrize("path", [p for p in R.golden_cases() if "cc_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
def test_cc_matches_reference_golden(ctx, path):
    z = R.load_golden(path)
    g = _dev_graph(ctx, z["off"], z["ids"], bool(z["symmetric"])).partition(int(z["P"]))
    for plan in (0, 1, 0):
        g.reorder(plan)
        lab, _ = gg.components(g)
        assert np.array_equal(lab.to_host(), z["cc_edge"].astype(np.int64)), plan
    g.close()


# ------------------------------------------------------------------ SpMV --

def test_spmv_matches_reference_golden(ctx):
    z = R.load_golden(R.GOLDEN + "/rmat10_dir_p16.npz")
    w = z["spmv_w"].astype(np.float32)
    g = _dev_graph(ctx, z["off"], z["ids"], w=w).partition(16)
    x = gg.DeviceArray.from_host(ctx, z["spmv_x"].astype(np.float64))
    y, _ = gg.spmv(g, x)
    assert R.rel_l1(y.to_host(), z["spmv_edge"]) < TOL64
    g.close()


@pytest.mark.parametrize("scale", [12, 18])
def test_spmv_float_weights(ctx, scale):
    g = gg.Graph.rmat(ctx, scale, 16, 3, weighted=True)
    off, ids, w = g.csr()
    x = oracle.uniform_pm1(3, g.n) if scale <= 12 else \
        np.random.default_rng(3).uniform(-1, 1, g.n)
    want = R.to_scipy(off, ids, w.astype(np.float64)).T @ x
    y64, _ = gg.spmv(g, gg.DeviceArray.from_host(ctx, x))
    assert R.rel_l1(y64.to_host(), want) < TOL64
    y32, _ = gg.spmv(g, gg.DeviceArray.from_host(ctx, x.astype(np.float32)))
    assert R.rel_l1(y32.to_host(), want) < TOL32
    g.close()


# ----------------------------------------------------- single edgeMaps ---

def _once(og, op, dense=None, sparse=None, bit=False, thr=-1, af=None, bf=None, ai=None, bi=None):
    n = og.n
    out = np.zeros(n, np.uint8)
    dm = oracle.C.c_int64()
    oe = oracle.C.c_int64()
    path = oracle.C.c_int()
    sp = None if sparse is None else np.ascontiguousarray(sparse, np.uint32)
    rc = oracle.lib().ggo_edgemap_once(
        og.h, op, None if dense is None else np.ascontiguousarray(dense, np.uint8).ctypes.data,
        None if sp is None else sp.ctypes.data, 0 if sp is None else len(sp), int(bit), thr,
        None if af is None else af.ctypes.data, None if bf is None else bf.ctypes.data,
        None if ai is None else ai.ctypes.data, None if bi is None else bi.ctypes.data, out,
        oracle.C.byref(dm), oracle.C.byref(oe), oracle.C.byref(path))
    assert rc == 0
    return out, dm.value, oe.value, path.value


@pytest.mark.parametrize("frac", [0.0005, 0.02, 0.3, 0.9])
def test_edgemap_cc_each_path(ctx, frac):
    off, ids = oracle.rmat_csr(14, 16, 2, symmetric=True)
    n = len(off) - 1
    rng = np.random.default_rng(int(frac * 1e4))
    dense = (rng.random(n) < frac).astype(np.uint8)
    ids0 = rng.permutation(n).astype(np.int64)
    og = oracle.Graph(off, ids, symmetric=True, P=64)
    a = ids0.copy()
    b = ids0.copy()  # prevIDs == IDs at round start
    want_f, dm, oe, path = _once(og, oracle.OP_CC, dense=dense, ai=a, bi=b)
    g = _dev_graph(ctx, off, ids, True).partition(64)
    F = gg.Frontier.from_dense(g, dense)
    A = gg.DeviceArray.from_host(ctx, ids0.astype(np.uint32))
    B = gg.DeviceArray.from_host(ctx, ids0.astype(np.uint32))
    out, gpath = gg.edge_map(g, F, gg.OP_CC, A, B)
    assert gpath == path
    got = out.dense()
    # one Jacobi round: labels and changed set are exact
    jac = ids0.copy()
    src = np.repeat(np.arange(n), np.diff(off))
    act = dense[src].astype(bool)
    np.minimum.at(jac, ids[act], ids0[src[act]])
    assert np.array_equal(A.to_host().astype(np.int64), jac)
    assert np.array_equal(got, (jac < ids0).astype(np.uint8))
    info = out.info()
    assert info["d_m"] == int(got.sum())
    assert info["out_edges"] == int(np.diff(off)[got.astype(bool)].sum())
    g.close()


@pytest.mark.parametrize("frac", [0.0005, 0.02, 0.3, 1.0])
def test_edgemap_pagerank_each_path(ctx, frac):
    off, ids = oracle.rmat_csr(14, 16, 1)
    n = len(off) - 1
    rng = np.random.default_rng(7)
    dense = (rng.random(n) < frac).astype(np.uint8)
    pc = rng.random(n)
    pn = rng.random(n)
    og = oracle.Graph(off, ids, P=64)
    a = pc.copy()
    b = pn.copy()
    want_f, dm, oe, path = _once(og, oracle.OP_PR, dense=dense, af=a, bf=b, bit=frac == 1.0)
    g = _dev_graph(ctx, off, ids).partition(64)
    F = gg.Frontier.bits(g) if frac == 1.0 else gg.Frontier.from_dense(g, dense)
    A = gg.DeviceArray.from_host(ctx, pc)
    B = gg.DeviceArray.from_host(ctx, pn)
    out, gpath = gg.edge_map(g, F, gg.OP_PAGERANK, A, B)
    assert gpath == path
    assert R.rel_l1(B.to_host(), b) < TOL64
    got = out.dense()
    assert np.array_equal(got, want_f)
    info = out.info()
    if path != oracle.PATH_SPARSE:  # sparse output lists carry duplicates
        assert info["d_m"] == dm
    assert info["out_edges"] == oe
    g.close()


@pytest.mark.parametrize("frac", [0.0002, 0.01, 0.2])
def test_edgemap_bfs_each_path(ctx, frac):
    off, ids = oracle.rmat_csr(14, 16, 2, symmetric=True)
    n = len(off) - 1
    rng = np.random.default_rng(11)
    front = rng.random(n) < frac
    par0 = np.where(front, np.arange(n), -1).astype(np.int64)
    og = oracle.Graph(off, ids, symmetric=True, P=64)
    a = par0.copy()
    want_f, dm, oe, path = _once(og, oracle.OP_BFS, dense=front.astype(np.uint8), ai=a)
    g = _dev_graph(ctx, off, ids, True).partition(64)
    F = gg.Frontier.from_dense(g, front.astype(np.uint8))
    P = gg.DeviceArray.from_host(ctx, par0.astype(np.int32))
    out, gpath = gg.edge_map(g, F, gg.OP_BFS, P)
    assert gpath == path
    got = out.dense()
    assert np.array_equal(got, want_f)  # newly reached set is exact
    par = P.to_host().astype(np.int64)
    assert np.array_equal(par >= 0, a >= 0)
    new = np.where(got)[0]
    assert np.all(front[par[new]])  # every new parent is in the frontier
    assert out.info()["d_m"] == dm and out.info()["out_edges"] == oe
    g.close()


def test_frontier_conversions(ctx):
    off, ids = oracle.rmat_csr(12, 8, 1)
    n = len(off) - 1
    g = _dev_graph(ctx, off, ids)
    rng = np.random.default_rng(1)
    flags = (rng.random(n) < 0.3).astype(np.uint8)
    F = gg.Frontier.from_dense(g, flags)
    assert np.array_equal(F.sparse(), np.where(flags)[0])  # packIndex order
    info = F.info()
    assert info["d_m"] == flags.sum()
    assert info["out_edges"] == np.diff(off)[flags.astype(bool)].sum()
    S = gg.Frontier.from_sparse(g, np.where(flags)[0][::-1])
    assert np.array_equal(S.dense(), flags)
    g.close()


def test_errors_are_reported(ctx):
    off, ids = oracle.rmat_csr(8, 8, 1)
    g = _dev_graph(ctx, off, ids)
    off2, ids2 = oracle.rmat_csr(9, 8, 1)
    g2 = _dev_graph(ctx, off2, ids2)
    F = gg.Frontier.bits(g2)
    A = gg.DeviceArray(ctx, g.n, np.float64).fill(0)
    with pytest.raises(gg.GGError) as e:
        gg.edge_map(g, F, gg.OP_PAGERANK, A, A)
    assert e.value.code == -4  # "Sizes Don't match" (ligra-numa.h:937-942)
    with pytest.raises(gg.GGError):
        gg.bfs(g, g.n + 5)
    with pytest.raises(gg.GGError):
        gg.Graph.from_csr(ctx, np.array([0, 2, 1], np.uint32), np.array([1, 0], np.uint32))
    g.close()
    g2.close()
