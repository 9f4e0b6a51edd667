"""Degenerate and extreme inputs through the C ABI against the oracle:
graphs with no edges, a single vertex, self-loops and repeated edges, a hub
whose in-list is split over many work items, a long path (many levels and
rounds), and the partition counts at their limits (P = 1, P = n)."""
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


def _csr(n, src, dst):
    """CSR keeping self-loops and repeated edges, lists in the given order."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.uint32)
    order = np.argsort(src, kind="stable")
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off).astype(np.uint32), dst[order]


def _check_all(ctx, off, ids, sym, P, root=0):
    n = len(off) - 1
    og = oracle.Graph(off, ids, symmetric=sym, P=P)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=sym).partition(P)
    if len(ids):
        want, it, _ = og.pagerank(10)
        got, info = gg.pagerank(g, np.float64)
        assert info.iters == it
        assert np.abs(got.to_host() - want).sum() <= 1e-12 * max(np.abs(want).sum(), 1e-300)
    else:
        # with no edges the first edgeMap returns partitioned_vertices::empty()
        # (0 rows) and the second one stops on "Sizes Don't match"
        # (ligra-numa.h:937-942, an abort in the reference)
        with pytest.raises(gg.GGError):
            gg.pagerank(g, np.float64)
    opar, olv, _ = og.bfs(root)
    par, lv, binfo = gg.bfs(g, root)
    assert binfo.iters == olv
    assert np.array_equal(par.to_host() >= 0, opar >= 0)
    lab, _ = gg.components(g)
    wl, _, _ = og.cc()
    assert np.array_equal(lab.to_host(), wl)
    x = np.random.default_rng(1).random(n)
    y, _ = gg.spmv(g, gg.DeviceArray.from_host(ctx, x))
    assert np.allclose(y.to_host(), R.to_scipy(off, ids).T @ x, rtol=1e-12, atol=0)
    g.close()


def test_no_edges(ctx):
    off, ids = np.zeros(6, np.uint32), np.zeros(0, np.uint32)
    _check_all(ctx, off, ids, False, 3)
    g = gg.Graph.from_csr(ctx, off, ids)
    sc, lv, _ = gg.bc(g, 2)
    s = sc.to_host()
    assert s[2] == 0 and np.all(np.isnan(np.delete(s, 2))) and list(lv.to_host()) == [-1, -1, 0, -1, -1]
    d, _ = gg.sssp(g, 1)
    assert list(d.to_host()) == [1073741823, 0, 1073741823, 1073741823, 1073741823]
    new, starts = g.vebo_order(2)
    assert sorted(new) == list(range(5)) and starts[-1] == 5
    g.close()


def test_single_vertex(ctx):
    off, ids = np.zeros(2, np.uint32), np.zeros(0, np.uint32)
    _check_all(ctx, off, ids, False, 1)


def test_self_loops_and_repeated_edges(ctx):
    # 0->0, 0->1 twice, 1->2, 2->0, 2->2, 3->1 three times
    off, ids = _csr(4, [0, 0, 0, 1, 2, 2, 3, 3, 3], [0, 1, 1, 2, 0, 2, 1, 1, 1])
    _check_all(ctx, off, ids, False, 2, root=3)


def test_hub_in_list_over_many_work_items(ctx):
    """Vertex 0 has 20,000 in-edges (split over work items of 2048 edges
    and combined in order) and 20,000 out-edges."""
    n = 20001
    src = np.concatenate([np.arange(1, n), np.zeros(n - 1, np.int64)])
    dst = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, n)])
    off, ids = _csr(n, src, dst)
    _check_all(ctx, off, ids, False, 64)


_PATHS = {oracle.PATH_SPARSE: "sparse", oracle.PATH_CSC: "csc", oracle.PATH_COO: "coo"}


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("root", [0, 5])
def test_bfs_first_level_either_branch(ctx, plan, root):
    """Hub 0 with 20,000 out-edges.  From the hub, level 1 is dense (1 +
    20,000 out-edges over m/20): a pull over the start's root bitmap with no
    conversion before it.  From a leaf, level 1 is a push and level 2 dense
    (the hub's out-degree), its input bitmap scattered from the pushed list.
    Levels, parents and the branch trace against the oracle."""
    n = 20001
    src = np.concatenate([np.arange(1, n), np.zeros(n - 1, np.int64)])
    dst = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, n)])
    off, ids = _csr(n, src, dst)
    print(f"bfs hub graph root={root} plan={plan}", flush=True)
    _, lv, tr = oracle.Graph(off, ids, symmetric=False, P=64).bfs(root)
    g = gg.Graph.from_csr(ctx, off, ids).partition(64).reorder(plan)
    p, l, info = gg.bfs(g, root)
    assert np.array_equal(l.to_host(), R.check_parents(off, ids, root, p.to_host()))
    assert info.iters == lv
    assert info.path_list() == [_PATHS[t.path] for t in tr]
    assert (info.path_list()[0] == "sparse") == (root != 0)
    g.close()


def test_long_path(ctx):
    """A directed path of 3000 vertices: 2999 BFS levels, 3000 CC rounds."""
    n = 3000
    off, ids = _csr(n, np.arange(n - 1), np.arange(1, n))
    _check_all(ctx, off, ids, False, 16)
    g = gg.Graph.from_csr(ctx, off, ids)
    _, lv, info = gg.bfs(g, 0)
    assert np.array_equal(lv.to_host(), np.arange(n)) and info.iters == n
    g.close()


@pytest.mark.parametrize("P", [1, 1 << 10])
def test_partition_count_limits(ctx, P):
    off, ids = oracle.rmat_csr(10, 8, 4)
    _check_all(ctx, off, ids, False, P, root=1)
