"""Components.C's loop on the device (gg_cc): each round's edgeMap branch is
chosen on the device and the labels are checked against scipy's component
minima (refcheck.cc_min_labels) and the oracle.  The graphs drive every
change of branch: a sparse round after a dense one (the dense round's flags
packed into a list), a dense round after a sparse one (prevIDs copied over
the pushed list first), runs of rounds past the first batches, and vertex
counts that are not whole 16-flag groups."""
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


def _sym_csr(n, src, dst):
    return oracle.csr_from_edges(n, np.asarray(src, np.uint32), np.asarray(dst, np.uint32), symmetric=True)


def _path_then_hub(L=30, leaves=2000):
    """Vertex 0 starts a path of L further vertices (ids leaves+1 ..
    leaves+L) that ends at a hub (the last of them) holding leaves 1 ..
    leaves.  Label 0 walks the path in sparse rounds; the round that lowers
    the hub has a frontier of one vertex with leaves+1 out-edges, over
    m/20, so the next round is dense."""
    path = [0] + list(range(leaves + 1, leaves + L + 1))
    hub = path[-1]
    src = path[:-1] + [hub] * leaves
    dst = path[1:] + list(range(1, leaves + 1))
    n = leaves + L + 1
    return n, src, dst


@pytest.mark.parametrize("plan", [0, 1])
def test_cc_dense_after_sparse(ctx, plan):
    n, src, dst = _path_then_hub()
    off, ids = _sym_csr(n, src, dst)
    print(f"cc path+hub n={n} m={len(ids)} plan={plan}", flush=True)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).partition(16).reorder(plan)
    lab, info = gg.components(g)
    assert np.array_equal(lab.to_host(), np.zeros(n, np.int64))
    names = info.path_list()
    assert len(names) == info.iters and info.iters <= 64
    assert names[0] in ("coo", "csc")
    # some sparse round is followed by a dense one
    assert any(a == "sparse" and b in ("coo", "csc") for a, b in zip(names, names[1:])), names
    # and some dense round by a sparse one
    assert any(a in ("coo", "csc") and b == "sparse" for a, b in zip(names, names[1:])), names
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("n,edges,seed", [(1237, 1500, 1), (4099, 9000, 2), (20011, 30000, 3)])
def test_cc_random_components(ctx, n, edges, seed, plan):
    """Many components of every size, isolated vertices, n not a multiple
    of 16 (the flag pack's tail)."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, edges)
    dst = rng.integers(0, n, edges)
    off, ids = _sym_csr(n, src, dst)
    print(f"cc random n={n} m={len(ids)} seed={seed} plan={plan}", flush=True)
    want = R.cc_min_labels(off, ids)
    og = oracle.Graph(off, ids, symmetric=True, P=16)
    wo, _, _ = og.cc()
    assert np.array_equal(wo, want)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).partition(16).reorder(plan)
    for _ in range(2):  # the second call sizes its first batch from the first's rounds
        lab, info = gg.components(g)
        assert np.array_equal(lab.to_host(), want)
        assert info.iters >= 1
    g.close()


def test_cc_many_rounds(ctx):
    """A symmetric path whose ids are shuffled: label 0 walks it a hop or
    more per round, over many batches of enqueued rounds."""
    n = 5001
    rng = np.random.default_rng(7)
    order = rng.permutation(n)
    off, ids = _sym_csr(n, order[:-1], order[1:])
    print(f"cc shuffled path n={n}", flush=True)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=True).partition(8)
    lab, info = gg.components(g)
    assert np.array_equal(lab.to_host(), np.zeros(n, np.int64))
    assert info.iters > 16
    g.close()


def test_cc_directed(ctx):
    """A directed graph: labels flow along edges only (in-neighbours in the
    dense rounds, out-edges in the sparse ones), so each vertex ends at the
    least id that reaches it, as in the oracle's loop."""
    off, ids = oracle.rmat_csr(13, 8, 5, symmetric=False)
    og = oracle.Graph(off, ids, symmetric=False, P=32)
    want, _, _ = og.cc()
    g = gg.Graph.from_csr(ctx, off, ids).partition(32).reorder(1)
    lab, _ = gg.components(g)
    assert np.array_equal(lab.to_host(), want)
    g.close()
