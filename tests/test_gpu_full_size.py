"""GPU parity at the sizes the bench quotes its numbers on.

- PageRank.C semantics on directed R-MAT scale 24 (seed 1, 384 edge-balanced
  partitions, the hot-sources-first plan the bench uses): FP64 within 1e-12
  and FP32 within 1e-5 relative L1 of the OpenMP C oracle, 10 iterations,
  bitwise equal run to run.  The device's generated CSR equals the oracle
  generator's bit for bit.
- BFS from root 100 and Components on symmetric R-MAT scale 24 (seed 2):
  levels bit-exact against the oracle's BFS.C restatement, parents a valid
  BFS tree, component labels equal to scipy's connected components labelled
  by their minimum vertex.
- Config 1 (directed R-MAT scale 20, 384 partitions) against the reference
  engine itself: the ranks that oracle/_ref/ref_pagerank (the reference's
  PageRank.C and engine, compiled from its own sources) writes, within 1e-12.
- The batched leg's 512-graph arena: a seeded sample of 16 graphs' PPR ranks
  and adjoint gradients (FP32, 30 iterations, captured loop) against the
  FP64 oracle run graph by graph, within 1e-5.

The host-side oracle graphs use vertex partitions (no COO lists to sort): the
values do not depend on the partitioning beyond FP summation order, which the
tolerances cover.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.sparse import csgraph, csr_matrix

import graphgrind_amd as gg
import oracle
import refcheck as R

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)


def _levels_from_parents(parents, root):
    """Depth of each vertex in a BFS tree given by its parent array."""
    n = len(parents)
    lvl = np.full(n, -1, np.int64)
    lvl[root] = 0
    p = parents.astype(np.int64)
    todo = np.flatnonzero((p >= 0) & (np.arange(n) != root))
    while len(todo):
        pl = lvl[p[todo]]
        done = pl >= 0
        lvl[todo[done]] = pl[done] + 1
        if not done.any():
            raise AssertionError("parent array is not a tree rooted at the root")
        todo = todo[~done]
    return lvl


def _check_parents_sorted(off, ids, root, parents, lvl):
    """BFS tree validity against a CSR whose rows are sorted: the root is its
    own parent, each reached v has the edge parent -> v and level one less."""
    n = len(off) - 1
    parents = parents.astype(np.int64)
    reached = lvl >= 0
    assert np.array_equal(parents >= 0, reached), "reached sets differ"
    assert parents[root] == root
    v = np.flatnonzero(reached & (np.arange(n) != root))
    p = parents[v]
    assert np.all(lvl[p] == lvl[v] - 1), "parent level mismatch"
    # edge p -> v: binary search of v inside p's sorted row
    off64 = off.astype(np.int64)
    lo, hi = off64[p], off64[p + 1]
    for _ in range(40):
        active = lo < hi
        if not active.any():
            break
        mid = (lo + hi) >> 1
        mv = ids[np.minimum(mid, len(ids) - 1)].astype(np.int64)
        go_right = active & (mv < v)
        lo = np.where(go_right, mid + 1, lo)
        hi = np.where(active & ~go_right, mid, hi)
    found = (lo < off64[p + 1]) & (ids[np.minimum(lo, len(ids) - 1)].astype(np.int64) == v)
    assert found.all(), f"{int((~found).sum())} parent edges missing"


def test_pagerank_rmat24_fp64_fp32_vs_oracle(ctx):
    _say("pagerank R-MAT 24: host graph")
    off, ids = oracle.rmat_csr(24, 16, 1)
    _say("pagerank R-MAT 24: device graph (384 partitions, hot-sources-first plan)")
    g = gg.Graph.rmat(ctx, 24, 16, 1)
    goff, gids, _ = g.csr()
    assert g.m == len(ids) and np.array_equal(goff, off) and np.array_equal(gids, ids)
    del goff, gids
    g.partition(384)
    g.reorder(1)
    r64, info = gg.pagerank(g, np.float64)
    assert info.iters == 10
    got64 = r64.to_host()
    r64b, _ = gg.pagerank(g, np.float64, out=r64)
    assert np.array_equal(r64b.to_host(), got64), "FP64 ranks differ between two runs"
    r32, info32 = gg.pagerank(g, np.float32)
    assert info32.iters == 10
    got32 = r32.to_host()
    g.close()
    _say("pagerank R-MAT 24: oracle")
    og = oracle.Graph(off, ids, P=384, part_vertex=True)
    want, it, _ = og.pagerank(10)
    del og
    assert it == 10
    e64, e32 = R.rel_l1(got64, want), R.rel_l1(got32, want)
    _say(f"pagerank R-MAT 24: rel L1 FP64 {e64:.3e} FP32 {e32:.3e}")
    assert e64 < 1e-12 and e32 < 1e-5


@pytest.fixture(scope="module")
def sym24():
    _say("symmetric R-MAT 24: host graph")
    off, ids = oracle.rmat_csr(24, 16, 2, symmetric=True)
    return off, ids


def test_bfs_rmat24_sym_levels_and_parents(ctx, sym24):
    off, ids = sym24
    g = gg.Graph.rmat(ctx, 24, 16, 2, symmetric=True)
    goff, gids, _ = g.csr()
    assert g.m == len(ids) and np.array_equal(goff, off) and np.array_equal(gids, ids)
    del goff, gids
    g.partition(384)
    g.reorder(1)
    _say("BFS R-MAT 24 sym: device")
    par, lv, info = gg.bfs(g, 100)
    gpar, glv = par.to_host(), lv.to_host()
    par2, _, info2 = gg.bfs(g, 100, parents=par, levels=False)  # the bench's call: Parents only
    assert np.array_equal(par2.to_host() >= 0, gpar >= 0) and info2.iters == info.iters
    g.close()
    _say("BFS R-MAT 24 sym: oracle")
    og = oracle.Graph(off, ids, symmetric=True, P=384, part_vertex=True)
    opar, olevels, _ = og.bfs(100)
    del og
    want = _levels_from_parents(opar, 100)
    assert info.iters == olevels
    assert np.array_equal(glv.astype(np.int64), want), "levels differ"
    _check_parents_sorted(off, ids, 100, gpar, want)
    _check_parents_sorted(off, ids, 100, par2.to_host(), want)
    _say(f"BFS R-MAT 24 sym: {olevels} levels, {int((want >= 0).sum())} reached")


def test_components_rmat24_sym_min_labels(ctx, sym24):
    off, ids = sym24
    g = gg.Graph.rmat(ctx, 24, 16, 2, symmetric=True)
    g.partition(384)
    g.reorder(1)
    _say("Components R-MAT 24 sym: device")
    lab, info = gg.components(g)
    got = lab.to_host().astype(np.int64)
    g.close()
    _say("Components R-MAT 24 sym: scipy")
    n = len(off) - 1
    A = csr_matrix((np.ones(len(ids), np.int8), ids, off), shape=(n, n))
    ncomp, cl = csgraph.connected_components(A, directed=False)
    mins = np.full(ncomp, np.iinfo(np.int64).max)
    np.minimum.at(mins, cl, np.arange(n))
    want = mins[cl]
    assert np.array_equal(got, want)
    _say(f"Components R-MAT 24 sym: {ncomp} components, {info.iters} rounds")


REF_PAGERANK = os.path.join(oracle.REF_DIR, "ref_pagerank")


@pytest.mark.skipif(not os.path.exists(REF_PAGERANK), reason="reference engine not built (oracle/_ref)")
@pytest.mark.parametrize("plan", [False, True])
def test_config1_vs_compiled_reference(ctx, plan):
    """Directed R-MAT scale 20, 384 edge partitions, 10 iterations: the
    reference's own PageRank.C on its own engine writes the ranks."""
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "ranks.bin")
        env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "16"))
        r = subprocess.run([REF_PAGERANK, "-rmat", "20", "16", "1", "-c", "384", "-rounds", "1", "-o", out],
                           capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        raw = np.fromfile(out, np.float64)
    iters, want = int(raw[:1].view(np.int64)[0]), raw[1:]
    g = gg.Graph.rmat(ctx, 20, 16, 1).partition(384)
    if plan:
        g.reorder(1)
    got, info = gg.pagerank(g, np.float64)
    assert info.iters == iters == 10
    e = R.rel_l1(got.to_host(), want)
    _say(f"config 1 vs compiled reference (plan={plan}): rel L1 {e:.3e}")
    assert e < 1e-12
    g.close()


def _arena_specs(count=512, seed=5):
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 19, count)
    efs = rng.choice([8, 16, 32], count)
    return rng, scales, efs


def test_batched_arena_sample_vs_oracle(ctx):
    """The bench's 512-graph arena (scale 10..18, ef 8/16/32, seed 5): PPR with
    1-8 seeds per graph and its adjoint, FP32, exactly 30 iterations in the
    captured loop; 16 graphs picked by a seeded draw are checked one by one."""
    count, iters, alpha = 512, 30, 0.85
    rng, scales, efs = _arena_specs(count)
    parts = [gg.Graph.rmat(ctx, int(s), int(e), 5 * 100003 + i) for i, (s, e) in enumerate(zip(scales, efs))]
    arena, vb, eb = gg.Graph.arena(ctx, parts)
    ones = gg.DeviceArray(ctx, max(arena.m, 1), np.float32).fill(1.0)
    arena.set_weights_dev(ones)
    _say(f"arena: {count} graphs, n={arena.n} m={arena.m}")
    n = arena.n
    v = np.zeros(n, np.float32)
    for i in range(count):
        ni = int(vb[i + 1] - vb[i])
        k = int(rng.integers(1, 9))
        v[vb[i] + rng.choice(ni, k, replace=False)] += np.float32(1.0 / k)
    vd = gg.DeviceArray.from_host(ctx, v)
    gg.autotune_pull(arena, np.float32, trials=2)
    r, info = gg.wpagerank(arena, np.float32, alpha, -1.0, iters, v=vd)
    assert info.iters == iters
    rh = r.to_host()
    u = np.zeros(n, np.float32)
    for i in range(count):
        t = vb[i] + rng.choice(int(vb[i + 1] - vb[i]), 8, replace=False)
        u[t] = 2 * rh[t]
    ud = gg.DeviceArray.from_host(ctx, u)
    gw, gv, _, ainfo = gg.wpagerank_adjoint(arena, r, ud, alpha, -1.0, iters)
    assert ainfo.iters == iters
    gwh, gvh = gw.to_host()[: arena.m], gv.to_host()
    pick = np.sort(np.random.default_rng(55).choice(count, 16, replace=False))
    for i in pick:
        p = parts[i]
        po, pi, _ = p.csr()
        w1 = np.ones(len(pi), np.float32)
        sl, el = slice(int(vb[i]), int(vb[i + 1])), slice(int(eb[i]), int(eb[i + 1]))
        want, it, _ = oracle.wpagerank(po, pi, w1, alpha, 0.0, iters, v=v[sl].astype(np.float64))
        assert it == iters
        e_r = R.rel_l1(rh[sl], want)
        _, gw_w, gv_w, ait = oracle.wpagerank_adjoint(po, pi, w1, rh[sl].astype(np.float64),
                                                      u[sl].astype(np.float64), alpha, 0.0, iters)
        # targets that all hold rank 0 (none reachable from the graph's
        # seeds) give u = 0 on this graph: the oracle's residual is exactly 0
        # after one step and it stops there; the gradients are 0 either way
        assert ait == iters or not u[sl].any()
        if not u[sl].any():
            assert not gvh[sl].any() and not gwh[el].any()
            _say(f"arena graph {i}: u = 0 (targets unreached), zero gradients")
            continue
        e_v = R.rel_l1(gvh[sl], gv_w)
        e_w = R.rel_l1(gwh[el], gw_w[:len(pi)]) if len(pi) else 0.0
        _say(f"arena graph {i}: scale {scales[i]} ef {efs[i]} n={p.n} m={p.m}: "
             f"ranks {e_r:.2e} gv {e_v:.2e} gw {e_w:.2e}")
        assert e_r < 1e-5 and e_v < 1e-5 and e_w < 1e-4
    arena.close()
    for p in parts:
        p.close()
