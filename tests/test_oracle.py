"""Pins the C oracle before anything is checked against it.

Each oracle function is checked against something it does not share code
with: the reference engine's own outputs (golden vectors written by
oracle/_ref from the reference sources), textbook algorithms (BFS levels,
union-find components, scipy SpMV), closed forms (PageRank iteration 1) and
a dense re-statement of the frontier-gated PageRank recurrence.
"""
import numpy as np
import pytest

import oracle
import refcheck as R


# ------------------------------------------------------------- generator --

@pytest.mark.parametrize("scale,ef,seed", [(6, 4, 1), (9, 16, 3), (11, 8, 5)])
def test_rmat_matches_numpy_spec(scale, ef, seed):
    src, dst, nraw = oracle.rmat_raw(scale, ef, seed)
    s2, d2 = R.rmat_raw_np(scale, ef, seed)
    assert np.array_equal(src[:nraw], s2) and np.array_equal(dst[:nraw], d2)


def test_rmat_skew():
    # a=.57: P(src bit 0) = a+b = .76 per level, P(dst bit 0) = a+c = .76
    src, dst, nraw = oracle.rmat_raw(12, 16, 9)
    for arr in (src[:nraw], dst[:nraw]):
        top = (arr >> 11) & 1
        assert abs(1 - top.mean() - 0.76) < 0.01


@pytest.mark.parametrize("sym", [False, True])
def test_csr_build(sym):
    src, dst, nraw = oracle.rmat_raw(10, 16, 2)
    off, ids = oracle.rmat_csr(10, 16, 2, symmetric=sym)
    o2, i2 = R.csr_np(1 << 10, src[:nraw], dst[:nraw], sym)
    assert np.array_equal(off, o2) and np.array_equal(ids, i2)
    # no self loops, sorted unique neighbour lists
    rows = np.repeat(np.arange(1 << 10), np.diff(off))
    assert not np.any(rows == ids)
    key = rows.astype(np.int64) * (1 << 32) + ids
    assert np.all(np.diff(key) > 0)
    if sym:
        A = R.to_scipy(off, ids)
        assert (A != A.T).nnz == 0


def test_transpose_vs_scipy():
    off, ids = oracle.rmat_csr(11, 8, 4)
    toff, tids, perm = oracle.transpose(off, ids)
    At = R.to_scipy(off, ids).T.tocsr()
    At.sort_indices()
    assert np.array_equal(toff, At.indptr) and np.array_equal(tids, At.indices)
    # perm maps CSC position -> CSR position of the same edge
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    cols = np.repeat(np.arange(len(off) - 1), np.diff(toff))
    assert np.array_equal(rows[perm], tids) and np.array_equal(ids[perm], cols)


def test_weights_and_vectors():
    w = oracle.weights(3, 5000)
    assert w.min() > 0 and w.max() <= 1.0
    k = np.round(w.astype(np.float64) * 2**24)
    assert np.all(k == w.astype(np.float64) * 2**24)  # exact dyadic
    x = oracle.uniform_pm1(3, 5000)
    assert x.min() >= -1 and x.max() < 1 and abs(x.mean()) < 0.05


# --------------------------------------------------------- partitioning --

def test_partition_vertex_mode():
    rc, s = oracle.partition_sizes(np.ones(1000, np.uint32), 1000, 7, mode=1)
    assert rc == 0 and list(s) == [142] * 6 + [148]


def test_partition_edge_mode_hand_case():
    # avg = 12/3 = 4: [3,1] reaches avg -> cut; [2,2] reaches avg -> cut; [4]
    deg = np.array([3, 1, 2, 2, 4], np.uint32)
    rc, s = oracle.partition_sizes(deg, 12, 3, mode=0)
    assert rc == 0 and list(s) == [2, 2, 1]
    # look-ahead cut: 2 < avg but 2 + next(3) > 1.1*avg closes the partition
    # (avg = 10/3 = 3; the last partition then absorbs the remainder)
    deg = np.array([2, 3, 3, 1, 1], np.uint32)
    rc, s = oracle.partition_sizes(deg, 10, 3, mode=0)
    assert rc == 0 and list(s) == [1, 1, 3]


def test_partition_overrun_reported():
    # [3,1] | [2] (look-ahead) | [3] then the look-ahead fires again with no
    # guard: the reference indexes partition P (graph-numa.h:1335)
    rc, _ = oracle.partition_sizes(np.array([3, 1, 2, 3, 3], np.uint32), 12, 3, mode=0)
    assert rc == -2


def test_partition_underrun_reported():
    # one vertex holds most edges: the greedy uses fewer than P partitions
    deg = np.array([90, 1, 1, 1, 1, 1, 1, 1, 1, 2], np.uint32)
    rc, s = oracle.partition_sizes(deg, 100, 4, mode=0)
    assert rc == 1 and s.sum() == 10 and s[-1] == 0


def test_partition_relabel_mode():
    deg = np.array([5, 4, 4, 6, 2, 3, 1], np.uint32)
    rc, s = oracle.partition_sizes(deg, 25, 3, mode=0, relabel=1)
    assert rc == 0 and list(s) == [3, 2, 2]


# ------------------------------------------------------------------ hilbert --

def test_hilbert_small_cases():
    # the order-2 curve visits (0,0),(0,1),(1,1),(1,0)
    assert [oracle.hilbert_xy2d(2, x, y) for x, y in [(0, 0), (0, 1), (1, 1), (1, 0)]] == \
        [0, 1, 2, 3]


@pytest.mark.parametrize("side", [4, 16, 64])
def test_hilbert_bijection_and_adjacency(side):
    pts = {}
    for x in range(side):
        for y in range(side):
            pts[oracle.hilbert_xy2d(side, x, y)] = (x, y)
    assert sorted(pts) == list(range(side * side))
    for d in range(side * side - 1):  # consecutive cells are grid neighbours
        (x0, y0), (x1, y1) = pts[d], pts[d + 1]
        assert abs(x0 - x1) + abs(y0 - y1) == 1


def test_round_up_pow2():
    assert [oracle.lib().ggo_round_up_pow2(v) for v in [1, 2, 3, 5, 64, 65, 41891]] == \
        [1, 2, 4, 8, 64, 128, 65536]
    assert oracle.lib().ggo_round_up_pow2(0) == 0


# ------------------------------------------------- vs the reference engine --

GOLD = R.golden_cases()


def _graph(z, part_vertex):
    return oracle.Graph(z["off"], z["ids"], symmetric=bool(z["symmetric"]), P=int(z["P"]),
                        part_vertex=part_vertex)


@pytest.mark.parametrize("path", GOLD, ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("part_vertex", [False, True])
def test_partitions_match_reference(path, part_vertex):
    z = R.load_golden(path)
    tag = "vertex" if part_vertex else "edge"
    g = _graph(z, part_vertex)
    assert np.array_equal(g.starts(0), z[f"coo_starts_{tag}"])
    assert np.array_equal(g.starts(1), z[f"csc_starts_{tag}"])
    if part_vertex:
        return
    counts = z[f"coo_count_{tag}"]
    edges = z[f"coo_edges_{tag}"]
    o = 0
    for p in range(int(z["P"])):
        s, d = g.coo(p)
        ref = edges[o:o + counts[p]]
        o += counts[p]
        assert len(s) == counts[p]
        side = oracle.lib().ggo_round_up_pow2(int(counts[p]))
        k_or = [oracle.hilbert_xy2d(side, int(a), int(b)) for a, b in zip(s, d)]
        k_rf = [oracle.hilbert_xy2d(side, int(a), int(b)) for a, b in ref]
        assert k_or == k_rf  # same Hilbert key sequence (ties may permute)
        assert sorted(zip(s.tolist(), d.tolist())) == sorted(map(tuple, ref.tolist()))


@pytest.mark.parametrize("path", [p for p in GOLD if "pagerank_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("part_vertex", [False, True])
def test_pagerank_matches_reference(path, part_vertex):
    z = R.load_golden(path)
    tag = "vertex" if part_vertex else "edge"
    g = _graph(z, part_vertex)
    x, it, tr = g.pagerank(10)
    assert it == int(z[f"pagerank_{tag}_iters"])
    assert R.rel_l1(x, z[f"pagerank_{tag}"]) < 1e-14
    assert R.rel_l1(x, R.pagerank_frontier_np(z["off"], z["ids"])) < 1e-12


def test_pagerank_iteration1_closed_form():
    off, ids = oracle.rmat_csr(10, 16, 1)
    n = len(off) - 1
    g = oracle.Graph(off, ids, P=16)
    x, it, tr = g.pagerank(1)
    outdeg = np.diff(off).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(off))
    want = 0.15 / n + 0.85 * np.bincount(ids, (1.0 / n) / outdeg[src], minlength=n)
    assert np.allclose(x, want, rtol=1e-13, atol=0)
    assert tr[0].path == oracle.PATH_COO and tr[0].in_active == n


def test_pagerank_frontier_semantics_probe():
    # in-degree-0 vertices follow q,0,q,q,0,q,... (frontier-gated semantics)
    off, ids = oracle.rmat_csr(12, 16, 1)
    n = len(off) - 1
    indeg = np.bincount(ids, minlength=n)
    g = oracle.Graph(off, ids, P=64)
    q = 0.15 / n
    seq = []
    for k in range(1, 7):
        x, _, _ = g.pagerank(k)
        seq.append(x[np.argmin(indeg)])
    assert np.allclose(seq, [q, 0, q, q, 0, q], rtol=1e-12, atol=0)


@pytest.mark.parametrize("path", [p for p in GOLD if "bfs_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("part_vertex", [False, True])
def test_bfs_matches_reference_levels(path, part_vertex):
    z = R.load_golden(path)
    tag = "vertex" if part_vertex else "edge"
    root = int(z["root"])
    g = _graph(z, part_vertex)
    par, levels, tr = g.bfs(root)
    lvl = R.check_parents(z["off"], z["ids"], root, par)
    R.check_parents(z["off"], z["ids"], root, z[f"bfs_{tag}"].astype(np.int64), lvl)
    assert levels == int(z[f"bfs_{tag}_iters"])
    # Parents agree except where the COO path picked among edges with equal
    # Hilbert keys (std::sort leaves their order unspecified).
    same = par == z[f"bfs_{tag}"].astype(np.int64)
    if not any(t.path == oracle.PATH_COO for t in tr):
        assert same.all()


@pytest.mark.parametrize("path", [p for p in GOLD if "cc_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("part_vertex", [False, True])
def test_cc_matches_reference(path, part_vertex):
    z = R.load_golden(path)
    tag = "vertex" if part_vertex else "edge"
    g = _graph(z, part_vertex)
    ids, rounds, tr = g.cc()
    assert np.array_equal(ids, z[f"cc_{tag}"].astype(np.int64))
    assert rounds == int(z[f"cc_{tag}_iters"])
    if int(z["symmetric"]):
        assert np.array_equal(ids, R.cc_min_labels(z["off"], z["ids"]))


@pytest.mark.parametrize("path", [p for p in GOLD if "spmv_edge" in np.load(p).files],
                         ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("part_vertex", [False, True])
def test_spmv_matches_reference(path, part_vertex):
    z = R.load_golden(path)
    tag = "vertex" if part_vertex else "edge"
    w = z["spmv_w"].astype(np.float64)
    g = oracle.Graph(z["off"], z["ids"], w=w, P=int(z["P"]), part_vertex=part_vertex)
    x = z["spmv_x"] if "spmv_x" in z else None
    y, tr = g.spmv(x)
    assert R.rel_l1(y, z[f"spmv_{tag}"]) < 1e-14
    n = len(z["off"]) - 1
    xv = x if x is not None else np.full(n, 1.0 / n)
    want = R.to_scipy(z["off"], z["ids"], w).T @ xv
    assert R.rel_l1(y, want) < 1e-13


def test_spmv_float_weights_vs_scipy():
    off, ids = oracle.rmat_csr(10, 16, 3)
    w = oracle.weights(3, len(ids)).astype(np.float64)
    x = oracle.uniform_pm1(3, len(off) - 1)
    g = oracle.Graph(off, ids, w=w, P=16)
    y, _ = g.spmv(x)
    assert R.rel_l1(y, R.to_scipy(off, ids, w).T @ x) < 1e-13


def test_plus_reduce_blocked_order():
    a = np.random.default_rng(0).random(5000)
    assert abs(oracle.plus_reduce(a) - a.sum()) < 1e-12 * a.sum()
    # serial left-to-right inside each 1024-element block, then over blocks
    blk = [sum(a[i * 1024:(i + 1) * 1024].tolist()) for i in range(5)]
    assert oracle.plus_reduce(a) == sum(blk)


# ------------------------------------- further reference runs (ext_*.npz) --

def _ext_cases():
    return R.ext_golden_cases()


def _coo_lists(og, P):
    return [np.stack(og.coo(p), 1) for p in range(P)]


def _split(counts, edges):
    out, o = [], 0
    for c in counts:
        out.append(edges[o:o + int(c)])
        o += int(c)
    return out


@pytest.mark.parametrize("path", _ext_cases(), ids=lambda p: p.split("/")[-1])
def test_relabel_partitions_match_reference(path):
    """-o (partition_relabel, graph-numa.h:1180-1194, 1290-1305): partition
    starts exact.  The COO lists hold the same edges with the same Hilbert key
    sequence; the order of edges with equal keys is std::sort's (unstable)
    and not pinned: relabel's small partitions give curve sides below the
    ids, where keys fold together."""
    z = np.load(path)
    off, ids, sym, P = z["off"], z["ids"], bool(z["symmetric"]), int(z["P"])
    for pv in (False, True):
        tag = "vertex" if pv else "edge"
        og = oracle.Graph(off, ids, symmetric=sym, P=P, part_vertex=pv, relabel=True)
        assert np.array_equal(og.starts(0), z[f"coo_starts_{tag}_relabel"])
        assert np.array_equal(og.starts(1), z[f"csc_starts_{tag}_relabel"])
        if pv:
            continue
        for p, (mine, ref) in enumerate(zip(_coo_lists(og, P), _split(z["coo_count_edge_relabel"],
                                                                          z["coo_edges_edge_relabel"]))):
            side = oracle.lib().ggo_round_up_pow2(len(ref))
            key = lambda e: [oracle.hilbert_xy2d(side, int(a), int(b)) for a, b in e]  # noqa: E731
            assert key(mine) == key(ref), p
            assert sorted(map(tuple, mine.tolist())) == sorted(map(tuple, ref.tolist())), p


@pytest.mark.parametrize("path", _ext_cases(), ids=lambda p: p.split("/")[-1])
def test_csr_order_coo_matches_reference(path):
    """EDGES_HILBERT=0: the COO lists in the reference's CSR order
    (graph-numa.h:338-350, 491-494), edge for edge."""
    z = np.load(path)
    off, ids, sym, P = z["off"], z["ids"], bool(z["symmetric"]), int(z["P"])
    og = oracle.Graph(off, ids, symmetric=sym, P=P, hilbert=False)
    for p, (mine, ref) in enumerate(zip(_coo_lists(og, P), _split(z["coo_count_edge_csr"], z["coo_edges_edge_csr"]))):
        assert np.array_equal(mine, ref), p


@pytest.mark.parametrize("path", R.ext_golden_cases("prbit_edge"),
                         ids=lambda p: p.split("/")[-1])
def test_pagerankbit_matches_reference_and_power_iteration(path):
    """app/PageRankBit.C: the reference's ranks, and (with every vertex active
    from the second iteration on and x reset before each pass) textbook
    power iteration x' = 0.85 A^T D^-1 x + 0.15/n from x = 1/n with the
    same L1 stop, dangling mass dropped."""
    z = np.load(path)
    off, ids, P = z["off"], z["ids"], int(z["P"])
    for pv in (False, True):
        tag = "vertex" if pv else "edge"
        r, it, _ = oracle.Graph(off, ids, P=P, part_vertex=pv).pagerank(10, bit_frontiers=True)
        assert it == int(z[f"prbit_{tag}_iters"])
        assert R.rel_l1(r, z[f"prbit_{tag}"]) < 1e-12
    n = len(off) - 1
    A = R.to_scipy(off, ids).tocsr()
    od = np.diff(off.astype(np.int64))
    c = np.full(n, 1.0 / n)
    for _ in range(10):
        x = 0.85 * (A.T @ np.where(od > 0, c / np.maximum(od, 1), 0.0)) + 0.15 / n
        done = np.abs(c - x).sum() < 1e-7
        c = x
        if done:
            break
    assert R.rel_l1(c, z["prbit_edge"]) < 1e-12


@pytest.mark.parametrize("path", _ext_cases(), ids=lambda p: p.split("/")[-1])
def test_frontier_trace_matches_reference(path):
    """The frontier's d_m and num_out_edges before each edgeMap (the public
    fields of partitioned_vertices, ligra-numa.h:91-272), recorded inside
    the reference's own PageRank.C / BFS.C / Components.C loops."""
    z = np.load(path)
    off, ids, sym, P, root = z["off"], z["ids"], bool(z["symmetric"]), int(z["P"]), int(z["root"])
    for pv in (False, True):
        tag = "vertex" if pv else "edge"
        og = oracle.Graph(off, ids, symmetric=sym, P=P, part_vertex=pv)
        for app in ("pagerank", "bfs", "cc"):
            key = f"trace_{app}_{tag}"
            if key not in z.files:
                continue
            if app == "pagerank":
                _, it, tr = og.pagerank(10)
            elif app == "bfs":
                _, it, tr = og.bfs(root)
            else:
                _, it, tr = og.cc()
            assert it == int(z[key + "_iters"]), (app, tag)
            mine = np.array([[t.in_active, t.in_out_edges] for t in tr], np.int64).reshape(-1, 2)
            assert np.array_equal(mine, z[key]), (app, tag)
