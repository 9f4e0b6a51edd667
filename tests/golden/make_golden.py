"""Writes the golden vectors under tests/golden/ from the REFERENCE engine.

Runs oracle/_ref/ref_<app> (the reference's own edgeMap/vertexMap, functors
and partitioned_graph, compiled from /root/reference by oracle/Makefile) on
small graphs and stores inputs and outputs as .npz fixtures.  Run in the
build container, where /root/reference exists:

    make -C oracle && python tests/golden/make_golden.py        # all fixtures
    make -C oracle && python tests/golden/make_golden.py ext    # ext_*.npz only

Inputs: the reference's own 'sixnode' graph (its only shipped fixture, kept
here as data) and seeded R-MAT graphs from the shared generator.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

# sixnode (reference file 'sixnode'): AdjacencyGraph, n=6, m=15
SIX_OFF = np.array([0, 2, 4, 7, 10, 14, 15], np.uint32)
SIX_IDS = np.array([1, 2, 2, 4, 1, 4, 5, 0, 2, 4, 0, 1, 2, 5, 1], np.uint32)


def int_weights(seed, m):
    # integer edge weights 1..19 (the reference's SPMV takes intE weights)
    return np.array([1 + oracle.splitmix64(seed ^ e) % 19 for e in range(m)], np.int32)


def case(name, off, ids, symmetric, P, root, apps, wseed=None, x=None, store=None):
    out = {"off": off, "ids": ids, "symmetric": np.int64(symmetric), "P": np.int64(P),
           "root": np.int64(root)}
    for part_vertex in (False, True):
        tag = "vertex" if part_vertex else "edge"
        for app in apps:
            w = int_weights(wseed, len(ids)) if app == "spmv" else None
            res, iters, parts = oracle.run_ref(app, off, ids, w=w, symmetric=symmetric, P=P,
                                               part_vertex=part_vertex, root=root, x=x,
                                               with_partitions=True, threads=1)
            out[f"{app}_{tag}"] = res
            out[f"{app}_{tag}_iters"] = np.int64(iters)
            if app == "spmv":
                out["spmv_w"] = w
                if x is not None:
                    out["spmv_x"] = x
            coo_starts, csc_starts, lists = parts
            out[f"coo_starts_{tag}"] = coo_starts
            out[f"csc_starts_{tag}"] = csc_starts
            if lists:
                out[f"coo_count_{tag}"] = np.array([len(l) for l in lists], np.int64)
                out[f"coo_edges_{tag}"] = (np.concatenate(lists) if lists else
                                           np.zeros((0, 2), np.uint32))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def case_apps(name, off, ids, symmetric, P, root, apps, wseed=None, part_vertex=False):
    """Fixtures of the further apps (BC.C, BellmanFord.C, PageRankDelta.C,
    PageRankConverage.C) on one graph, single-threaded reference runs."""
    out = {"off": off, "ids": ids, "symmetric": np.int64(symmetric), "P": np.int64(P),
           "root": np.int64(root), "part_vertex": np.int64(part_vertex)}
    for app in apps:
        w = int_weights(wseed, len(ids)) if app == "bellmanford" else None
        res, iters, _ = oracle.run_ref(app, off, ids, w=w, symmetric=symmetric, P=P, root=root, threads=1,
                                       part_vertex=part_vertex)
        out[app] = res
        out[f"{app}_iters"] = np.int64(iters)
        if w is not None:
            out["bellmanford_w"] = w
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def case_ext(name, off, ids, symmetric, P, root, trace_apps):
    """Further reference runs on one graph (ext_<name>.npz): partition
    starts and COO lists with partition_relabel (the driver's -o), COO lists
    in the reference's CSR order (EDGES_HILBERT=0), PageRankBit.C's ranks,
    and the frontier's (d_m, num_out_edges) before every edgeMap of the
    listed apps."""
    out = {"off": off, "ids": ids, "symmetric": np.int64(symmetric), "P": np.int64(P),
           "root": np.int64(root)}
    for part_vertex in (False, True):
        tag = "vertex" if part_vertex else "edge"
        _, _, (cs, ss, lists) = oracle.run_ref("pagerank", off, ids, symmetric=symmetric, P=P,
                                               part_vertex=part_vertex, root=root, with_partitions=True,
                                               threads=1, relabel=True)
        out[f"coo_starts_{tag}_relabel"] = cs
        out[f"csc_starts_{tag}_relabel"] = ss
        if lists:
            out[f"coo_count_{tag}_relabel"] = np.array([len(l) for l in lists], np.int64)
            out[f"coo_edges_{tag}_relabel"] = np.concatenate(lists)
        if not symmetric:
            res, iters, _ = oracle.run_ref("prbit", off, ids, P=P, part_vertex=part_vertex, root=root, threads=1)
            out[f"prbit_{tag}"] = res
            out[f"prbit_{tag}_iters"] = np.int64(iters)
        for app in trace_apps:
            _, iters, _, tr = oracle.run_ref(app, off, ids, symmetric=symmetric, P=P, part_vertex=part_vertex,
                                             root=root, threads=1, trace=True)
            out[f"trace_{app}_{tag}"] = tr
            out[f"trace_{app}_{tag}_iters"] = np.int64(iters)
    _, _, (cs, ss, lists) = oracle.run_ref("pagerank_csr", off, ids, symmetric=symmetric, P=P, root=root,
                                           with_partitions=True, threads=1)
    out["coo_count_edge_csr"] = np.array([len(l) for l in lists], np.int64)
    out["coo_edges_edge_csr"] = np.concatenate(lists)
    path = os.path.join(HERE, f"ext_{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def main_ext():
    if not all(oracle.ref_available(a) for a in ("pagerank", "prbit", "pagerank_csr", "bfs", "cc")):
        sys.exit("oracle/_ref not built (needs /root/reference): make -C oracle")
    case_ext("sixnode_p2", SIX_OFF, SIX_IDS, False, 2, 0, ["pagerank", "bfs"])
    so, si = oracle.csr_from_edges(6, np.repeat(np.arange(6), np.diff(SIX_OFF)), SIX_IDS, symmetric=True)
    case_ext("sixnode_sym_p2", so, si, True, 2, 0, ["bfs", "cc"])
    off, ids = oracle.rmat_csr(10, 16, 1)
    case_ext("rmat10_dir_p16", off, ids, False, 16, 100, ["pagerank", "bfs"])
    off, ids = oracle.rmat_csr(11, 16, 2, symmetric=True)
    case_ext("rmat11_sym_p32", off, ids, True, 32, 100, ["bfs", "cc"])
    off, ids = oracle.rmat_csr(12, 8, 5)
    case_ext("rmat12_dir_p64", off, ids, False, 64, 100, ["pagerank", "bfs"])


def main():
    if not all(oracle.ref_available(a) for a in ("pagerank", "bc", "bellmanford", "prdelta", "prconv")):
        sys.exit("oracle/_ref not built (needs /root/reference): make -C oracle")
    apps_dir = ["pagerank", "bfs", "cc", "spmv"]
    case("sixnode_p2", SIX_OFF, SIX_IDS, False, 2, 0, apps_dir, wseed=7)
    # P=3 is not used: the greedy overruns its arrays on sixnode (UB in the reference)
    case("sixnode_p6", SIX_OFF, SIX_IDS, False, 6, 0, apps_dir, wseed=7)
    so, si = oracle.csr_from_edges(6, np.repeat(np.arange(6), np.diff(SIX_OFF)), SIX_IDS,
                                   symmetric=True)
    case("sixnode_sym_p2", so, si, True, 2, 0, ["bfs", "cc"])
    off, ids = oracle.rmat_csr(10, 16, 1)
    x = oracle.uniform_pm1(3, len(off) - 1)
    case("rmat10_dir_p16", off, ids, False, 16, 100, apps_dir, wseed=3, x=x)
    off, ids = oracle.rmat_csr(11, 16, 2, symmetric=True)
    case("rmat11_sym_p32", off, ids, True, 32, 100, ["bfs", "cc"])
    off, ids = oracle.rmat_csr(12, 8, 5)
    case("rmat12_dir_p64", off, ids, False, 64, 100, ["pagerank", "bfs"])
    # further apps: BC on symmetric graphs (its GA.transpose() is then the
    # identity in the reference too), the others on directed graphs
    case_apps("apps_sixnode_p2", SIX_OFF, SIX_IDS, False, 2, 0, ["bellmanford", "prdelta", "prconv"], wseed=7)
    case_apps("apps_sixnode_sym_p2", so, si, True, 2, 0, ["bc"])
    # BC on a directed graph, vertex partitions (no COO lists, whose
    # direction the reference's transpose() keeps): from root 100 the deepest
    # level has no out-edges, so its stored out-degree sum is 0 and the
    # reference's backward phase starts one level up
    d_off, d_ids = oracle.rmat_csr(10, 16, 1)
    case_apps("apps_rmat10_dir_vertex_p16", d_off, d_ids, False, 16, 100, ["bc"], part_vertex=True)
    off, ids = oracle.rmat_csr(10, 16, 2, symmetric=True)
    case_apps("apps_rmat10_sym_p16", off, ids, True, 16, 3, ["bc"])
    off, ids = oracle.rmat_csr(10, 16, 1)
    case_apps("apps_rmat10_dir_p16", off, ids, False, 16, 0, ["bellmanford", "prdelta", "prconv"], wseed=3)


if __name__ == "__main__":
    # 'ext' writes only the ext_*.npz fixtures (the others stay as committed)
    main_ext() if sys.argv[1:] == ["ext"] else main()
