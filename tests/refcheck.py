"""Independent (numpy/scipy) restatements used to pin the oracle and the GPU.

Nothing here calls the oracle: these are textbook or closed-form versions of
what the oracle and the HIP path must agree with.
"""
from __future__ import annotations

import glob
import os

import numpy as np
import scipy.sparse as sp
from scipy.sparse import csgraph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_np(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rmat_raw_np(scale, ef, seed):
    """Same R-MAT spec as the oracle, written independently in numpy."""
    ne = ef << scale
    e = np.arange(ne, dtype=np.uint64)
    src = np.zeros(ne, np.uint64)
    dst = np.zeros(ne, np.uint64)
    ta = 2448131359
    tab = ta + 816043786
    tabc = tab + 816043786
    for lvl in range(scale - 1, -1, -1):
        key = (np.uint64(seed) << np.uint64(40)) ^ (e << np.uint64(6)) ^ np.uint64(lvl)
        r = (splitmix64_np(key) >> np.uint64(32)).astype(np.int64)
        sb = (r >= tab).astype(np.uint64)
        db = (((r >= ta) & (r < tab)) | (r >= tabc)).astype(np.uint64)
        src |= sb << np.uint64(lvl)
        dst |= db << np.uint64(lvl)
    return src.astype(np.uint32), dst.astype(np.uint32)


def csr_np(n, src, dst, symmetric=False):
    keep = src != dst
    s = src[keep].astype(np.int64)
    d = dst[keep].astype(np.int64)
    if symmetric:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    key = np.unique(s * (1 << 32) + d)
    s = key >> 32
    d = key & 0xFFFFFFFF
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.uint32), d.astype(np.uint32)


def to_scipy(off, ids, w=None):
    n = len(off) - 1
    data = np.ones(len(ids)) if w is None else np.asarray(w, np.float64)
    return sp.csr_matrix((data, np.asarray(ids, np.int64), np.asarray(off, np.int64)),
                         shape=(n, n))


def expand_ranges(starts, ends):
    """Concatenation of arange(starts[i], ends[i]) without a Python loop."""
    starts = np.asarray(starts, np.int64)
    lens = np.asarray(ends, np.int64) - starts
    tot = int(lens.sum())
    if tot == 0:
        return np.zeros(0, np.int64)
    base = np.repeat(starts - (np.cumsum(lens) - lens), lens)
    return base + np.arange(tot, dtype=np.int64)


def bfs_levels(off, ids, root):
    """Textbook BFS hop distances (-1 unreached)."""
    n = len(off) - 1
    lvl = np.full(n, -1, np.int64)
    lvl[root] = 0
    frontier = np.array([root], np.int64)
    off = np.asarray(off, np.int64)
    ids = np.asarray(ids, np.int64)
    d = 0
    while len(frontier):
        d += 1
        nb = np.unique(ids[expand_ranges(off[frontier], off[frontier + 1])])
        nb = nb[lvl[nb] < 0]
        lvl[nb] = d
        frontier = nb
    return lvl


def min_parents(off, ids, lvl):
    """parent[v] = smallest u with lvl[u] = lvl[v]-1 and u->v (root: itself)."""
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(np.asarray(off, np.int64)))
    dst = np.asarray(ids, np.int64)
    ok = (lvl[src] >= 0) & (lvl[dst] == lvl[src] + 1)
    par = np.full(n, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(par, dst[ok], src[ok])
    par[lvl < 0] = -1
    par[lvl == 0] = np.where(lvl == 0)[0]
    return par


def check_parents(off, ids, root, parents, lvl=None):
    """Parent validity: root is its own parent, each reached v has an edge
    parent->v with level one less; unreached exactly where lvl < 0."""
    if lvl is None:
        lvl = bfs_levels(off, ids, root)
    parents = np.asarray(parents, np.int64)
    reached = lvl >= 0
    assert np.array_equal(parents >= 0, reached), "reached sets differ"
    assert parents[root] == root
    v = np.where(reached & (np.arange(len(lvl)) != root))[0]
    p = parents[v]
    assert np.all(lvl[p] == lvl[v] - 1), "parent level mismatch"
    A = to_scipy(off, ids).tocsr()
    assert np.all(np.asarray(A[p, v]).ravel() > 0), "parent edge missing"
    return lvl


def cc_min_labels(off, ids):
    """Weakly connected components labelled by their minimum vertex id."""
    A = to_scipy(off, ids)
    ncomp, lab = csgraph.connected_components(A, directed=True, connection="weak")
    mins = np.full(ncomp, np.iinfo(np.int64).max)
    np.minimum.at(mins, lab, np.arange(len(lab)))
    return mins[lab]


def pagerank_frontier_np(off, ids, iters=10, damping=0.85, eps=1e-7):
    """The frontier-gated recurrence of PageRank.C (x accumulates onto stale
    values, vertexMap/reset only over the input frontier) in dense numpy."""
    n = len(off) - 1
    off = np.asarray(off, np.int64)
    outdeg = np.diff(off)
    src = np.repeat(np.arange(n), outdeg)
    dst = np.asarray(ids, np.int64)
    c = np.full(n, 1.0 / n)
    x = np.zeros(n)
    F = np.ones(n, bool)
    for _ in range(iters):
        act = F[src]
        x = x + np.bincount(dst[act], (c[src] / outdeg[src])[act], minlength=n)
        nxt = np.zeros(n, bool)
        nxt[dst[act]] = True
        x[F] = damping * x[F] + (1 - damping) * (1.0 / n)
        c = np.abs(c - x)
        if c.sum() < eps:
            return x
        c[F] = 0.0
        c, x = x, c
        F = nxt
    return c


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def golden_cases():
    """Fixtures of the engine apps (partitions, COO lists, the four apps)."""
    return sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("apps_", "ext_")))


def ext_golden_cases(key=None):
    """Further reference runs (ext_*.npz: -o relabel partitions, CSR-order
    COO lists, PageRankBit.C, per-step frontier traces), optionally only
    those holding 'key'."""
    paths = sorted(glob.glob(os.path.join(GOLDEN, "ext_*.npz")))
    if key is None:
        return paths
    return [p for p in paths if key in np.load(p, allow_pickle=False).files]


def app_golden_cases(app=None):
    """Fixtures of the further apps (BC, BellmanFord, PageRankDelta,
    PageRankConverage), optionally only those holding 'app'."""
    paths = sorted(glob.glob(os.path.join(GOLDEN, "apps_*.npz")))
    if app is None:
        return paths
    return [p for p in paths if app in np.load(p, allow_pickle=False).files]


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}
