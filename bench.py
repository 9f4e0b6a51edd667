"""Benchmark: GraphGrind edgeMap on one MI355X.

Headline (the JSON 'value'): PageRank.C semantics on a directed R-MAT scale-24
graph (edge factor 16, a=.57 b=.19 c=.19 d=.05, seed 1, self loops and
duplicates removed, uint32 ids/offsets), 384 edge-balanced Hilbert COO
partitions, FP64 ranks, exactly 10 iterations.  One step = one Compute()
(initialisation + 10 iterations); GTEPS = 10*m / step time.  Inputs are
resident in HBM before the timed region (graph build and partitioning are
untimed, as in ligra-numa.h:1155-1157).

Also measured on the same line: BFS from root 100 and connected components
on the symmetric R-MAT scale-24 graph (seed 2), the FP32 PageRank variant,
weighted SpMV and its adjoint (FP32, R-MAT scale 24, seed 3), the
differentiable PageRank forward and implicit adjoint with the adjoint /
forward time ratio (FP64, R-MAT scale 22, seed 4, 4096 targets), the
batched personalised PageRank + adjoint over 512 packed R-MAT graphs and
32-source BFS on symmetric R-MAT scale 22, the roofline of the dominant
kernel (HIP events on the library stream) and the reference engine's CPU
time on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 24] [--no-extras] [--no-batch]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pagerank_bench(gg, ctx, scale, steps, warmup, dtype, P=384):
    t0 = time.perf_counter()
    g = gg.Graph.rmat(ctx, scale, 16, 1)
    g.partition(P)
    g.reorder(1)  # hot-sources-first numbering for the pull (untimed preprocessing)
    log(f"[bench] pagerank graph n={g.n} m={g.m} built+partitioned in {time.perf_counter() - t0:.1f}s")
    out = gg.DeviceArray(ctx, g.n, dtype)
    # the first call on the graph allocates the loop's rank, contribution and
    # frontier buffers (kept for later calls): the "including per-call
    # allocation and initialisation" figure
    ctx.sync()
    tf = time.perf_counter()
    gg.pagerank(g, dtype, out=out)
    ctx.sync()
    first_ms = (time.perf_counter() - tf) * 1e3
    for _ in range(warmup):
        gg.pagerank(g, dtype, out=out)
    ctx.sync()
    kern_ms, ev_ms = [], []
    t1 = time.perf_counter()
    for _ in range(steps):
        _, info = gg.pagerank(g, dtype, out=out)
        ev_ms.append(info.ms)
        ms, launches = ctx.kernel_time()
        kern_ms.append(ms)
        assert info.iters == 10 and launches == 10
    ctx.sync()
    wall = (time.perf_counter() - t1) / steps
    n, m = g.n, g.m
    item = 8 if dtype == np.float64 else 4
    # edge-centric no-reuse byte model per iteration: per edge a 4 B source id
    # and one gathered contribution; per vertex offset, contribution write,
    # rank read/write, out-degree and frontier flags
    algo_bytes = (4 + item) * m + (32 if item == 8 else 18) * n
    avg_kernel = float(np.mean(kern_ms)) / 1e3  # seconds per iteration
    vname = "double" if dtype == np.float64 else "float"
    xseg = g.get_option(gg.OPT_XCD_SEGMENTS)
    if xseg == 2 or (xseg == 1 and m >= (1 << 24)):
        # the source-segmented pull: partial pass, long-row fix-up and the
        # combine with the fused epilogue, timed together per iteration
        kern = f"k_xseg_part<{vname}> + k_rows_fix + k_xseg_combine<OpPRFused<{vname}>>"
    else:
        kern = ("k_rows" if g.get_option(gg.OPT_PULL_SCHEDULE) == 1 else "k_pull") + f"<OpPRFused<{vname}>>"
    res = {
        "kernel": kern,
        "n": n, "m": m, "step_ms": wall * 1e3, "event_ms": float(np.mean(ev_ms)),
        "gteps": 10 * m / wall / 1e9, "first_call_ms": first_ms,
        "median_event_ms": float(np.median(ev_ms)),
        "kernel_ms": avg_kernel * 1e3,
        "algo_bytes_per_launch": algo_bytes,
        "achieved_gbs": algo_bytes / avg_kernel / 1e9,
    }
    g.close()
    out.free()
    return res


def bfs_cc_bench(gg, ctx, scale, steps, root=100, P=384):
    t0 = time.perf_counter()
    g = gg.Graph.rmat(ctx, scale, 16, 2, symmetric=True)
    g.partition(P)
    g.reorder(1)  # BFS and Components run in the hot-sources-first numbering (untimed preprocessing)
    log(f"[bench] bfs graph n={g.n} m={g.m} built+partitioned in {time.perf_counter() - t0:.1f}s")
    # BFS.C computes Parents only (app/BFS.C:77-101); the level array is an
    # extra output of gg_bfs, left out of the timed calls
    par = gg.DeviceArray(ctx, g.n, np.int32)
    gg.bfs(g, root, parents=par, levels=False)  # warm-up
    ctx.sync()
    times, infos = [], []
    for _ in range(steps):
        t1 = time.perf_counter()
        _, _, info = gg.bfs(g, root, parents=par, levels=False)
        times.append(time.perf_counter() - t1)
        infos.append(info)
    deg = g.degrees(False)
    reached = par.to_host() >= 0
    # Graph500 convention: undirected edges inside the reached component
    und_edges = int(deg[reached].astype(np.int64).sum()) // 2
    t = float(np.median(times))
    info = infos[-1]
    res = {"root": root, "levels": info.iters, "paths": info.path_list(), "reached": int(reached.sum()),
           "edges_in_component": und_edges, "ms": t * 1e3, "gteps": und_edges / t / 1e9,
           "edges_scanned": int(info.edges_scanned), "plan": "hot-sources-first",
           "achieved_gbs": (4 * info.edges_scanned + 8 * g.n) / t / 1e9}
    lab = gg.DeviceArray(ctx, g.n, np.uint32)
    gg.components(g, out=lab)
    ctimes = []
    for _ in range(max(1, steps)):
        t1 = time.perf_counter()
        _, cinfo = gg.components(g, out=lab)
        ctimes.append(time.perf_counter() - t1)
    res["cc"] = {"rounds": cinfo.iters, "ms": float(np.median(ctimes)) * 1e3,
                 "paths": cinfo.path_list()[:16]}
    g.close()
    return res


def spmv_adjoint_bench(gg, ctx, scale, steps, seed=3, plan=True):
    """Weighted SpMV and its adjoint (FP32): y = A x over the in-edge twin;
    for L = 0.5|y - t|^2, dL/dx over the out-edge CSR and dL/dw per edge."""
    g = gg.Graph.rmat(ctx, scale, 16, seed, weighted=True)
    if plan:
        g.reorder(1)  # hot-sources-first gathers (untimed preprocessing, same sums)
    n, m = g.n, g.m
    rng = np.random.default_rng(seed)
    x = gg.DeviceArray.from_host(ctx, rng.uniform(-1, 1, n).astype(np.float32))
    t = rng.standard_normal(n).astype(np.float32)
    y = gg.DeviceArray(ctx, n, np.float32)
    gx = gg.DeviceArray(ctx, n, np.float32)
    gw = gg.DeviceArray(ctx, m, np.float32)
    gg.spmv(g, x, y)
    gy = gg.DeviceArray.from_host(ctx, y.to_host() - t)
    gg.spmv_bwd(g, x, gy, gx=gx, gw=gw)  # warm-up (builds the CSR schedule)
    fwd, bwd = [], []
    for _ in range(max(steps, 3)):
        _, i1 = gg.spmv(g, x, y)
        fwd.append(i1.ms)
        _, _, i2 = gg.spmv_bwd(g, x, gy, gx=gx, gw=gw)
        bwd.append(i2.ms)
    f, b = float(np.median(fwd)), float(np.median(bwd))
    # FP32: per edge 4 B id + 4 B weight + 4 B gathered x; per row offset + y
    fbytes = 12 * m + 8 * n
    # one fused pass over the out-edge CSR: per edge 4 B id + 4 B weight +
    # 4 B gathered gy + 4 B written gw; per row offset + x + gx
    bbytes = 16 * m + 12 * n
    res = {"n": n, "m": m, "fwd_ms": f, "bwd_ms": b, "bwd_fwd_ratio": b / f,
           "fwd_gbs": fbytes / f / 1e6, "bwd_gbs": bbytes / b / 1e6}
    g.close()
    return res


def wpagerank_adjoint_bench(gg, ctx, scale, steps, seed=4, targets=4096, alpha=0.85, eps=1e-7, plan=True):
    """Differentiable PageRank (FP64): forward to L1 step < 1e-7 (at most 100
    iterations), loss = sum over 'targets' vertices of r^2, implicit adjoint
    giving dL/dw (per edge) and dL/dv (teleport)."""
    g = gg.Graph.rmat(ctx, scale, 16, seed, weighted=True)
    if plan:
        g.reorder(1)  # both solves in the hot-sources-first numbering (untimed preprocessing)
    n, m = g.n, g.m
    r = gg.DeviceArray(ctx, n, np.float64)
    gw = gg.DeviceArray(ctx, m, np.float64)
    gv = gg.DeviceArray(ctx, n, np.float64)
    gg.wpagerank(g, np.float64, alpha, eps, 100, out=r)  # warm-up (weight sums, schedules)
    rh = r.to_host()
    T = np.random.default_rng(seed).choice(n, targets, replace=False)
    u = np.zeros(n)
    u[T] = 2 * rh[T]
    ud = gg.DeviceArray.from_host(ctx, u)
    gg.wpagerank_adjoint(g, r, ud, alpha, eps, 100, gw=gw, gv=gv)
    fwd, adj, fi, ai, fk, ak = [], [], [], [], [], []
    for _ in range(max(steps, 3)):
        _, i1 = gg.wpagerank(g, np.float64, alpha, eps, 100, out=r)
        fwd.append(i1.ms)
        fi.append(i1.iters)
        fk.append(ctx.kernel_time()[0])  # mean pull time per iteration
        _, _, _, i2 = gg.wpagerank_adjoint(g, r, ud, alpha, eps, 100, gw=gw, gv=gv)
        adj.append(i2.ms)
        ai.append(i2.iters)
        ak.append(ctx.kernel_time()[0])
    f, a = float(np.median(fwd)), float(np.median(adj))
    # per iteration, FP64: 4 B id + 4 B weight + 8 B gathered value per edge;
    # per row offset, old value, teleport/seed, weight sum, new value and
    # (forward) the next contribution
    it_bytes = 16 * m + 36 * n
    res = {"n": n, "m": m, "fwd_ms": f, "adj_ms": a, "adjoint_forward_ratio": a / f,
           "fwd_iters": fi[-1], "adj_iters": ai[-1],
           "fwd_iter_ms": float(np.median(fk)), "adj_iter_ms": float(np.median(ak)),
           "fwd_iter_gbs": it_bytes / float(np.median(fk)) / 1e6,
           "adj_iter_gbs": it_bytes / float(np.median(ak)) / 1e6}
    g.close()
    return res


def batched_bench(gg, ctx, count=512, seed=5, iters=30, msbfs_scale=22, plan=True):
    """Batched uneven workloads: 'count' directed R-MAT graphs (scale uniform
    in 10..18, edge factor in {8, 16, 32}, seeds derived from 'seed') packed
    into one arena; personalised PageRank with 1-8 seeds per graph (FP32,
    exactly 'iters' iterations) and its adjoint (loss: sum of r^2 over 8
    targets per graph), each one captured hipGraph replay, after the
    autotuner picked the pull schedule; beside it the same solves graph by
    graph, and as sub-batches of 64 graphs; the request latencies of the
    three under one serving model (all requests at t = 0, served back to
    back).  Then 32-source BFS on symmetric R-MAT scale 22 against 32
    single-source BFS runs, and BC.C's forward phase for the same 32
    sources in one pass beside 32 single-source BC.C runs."""
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 19, count)
    efs = rng.choice([8, 16, 32], count)
    t0 = time.perf_counter()
    parts = [gg.Graph.rmat(ctx, int(s), int(e), seed * 100003 + i) for i, (s, e) in enumerate(zip(scales, efs))]
    arena, vb, eb = gg.Graph.arena(ctx, parts)
    ones = gg.DeviceArray(ctx, max(arena.m, 1), np.float32).fill(1.0)
    arena.set_weights_dev(ones)
    log(f"[bench] batch: {count} graphs, n={arena.n} m={arena.m} built in {time.perf_counter() - t0:.1f}s")
    n = arena.n
    v = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    for i in range(count):
        ni = int(vb[i + 1] - vb[i])
        k = int(rng.integers(1, 9))
        v[vb[i] + rng.choice(ni, k, replace=False)] += np.float32(1.0 / k)
    vd = gg.DeviceArray.from_host(ctx, v)
    # no renumbered plan here: a global out-degree order interleaves the
    # packed graphs, and the batched solves ran 1.3x slower with it
    tune = gg.autotune_pull(arena, np.float32, trials=3)
    r = gg.DeviceArray(ctx, n, np.float32)
    gg.wpagerank(arena, np.float32, 0.85, -1.0, iters, v=vd, out=r)  # capture
    rh = r.to_host()
    for i in range(count):
        t = vb[i] + rng.choice(int(vb[i + 1] - vb[i]), 8, replace=False)
        u[t] = 2 * rh[t]
    ud = gg.DeviceArray.from_host(ctx, u)
    gw = gg.DeviceArray(ctx, max(arena.m, 1), np.float32)
    gv = gg.DeviceArray(ctx, n, np.float32)
    gg.wpagerank_adjoint(arena, r, ud, 0.85, -1.0, iters, gw=gw, gv=gv)  # capture
    fw, ad = [], []
    for _ in range(3):
        _, i1 = gg.wpagerank(arena, np.float32, 0.85, -1.0, iters, v=vd, out=r)
        fw.append(i1.ms)
        _, _, _, i2 = gg.wpagerank_adjoint(arena, r, ud, 0.85, -1.0, iters, gw=gw, gv=gv)
        ad.append(i2.ms)
    tb_f, tb_a = float(np.median(fw)), float(np.median(ad))
    # the same solves graph by graph (each its own captured loop)
    per_f, per_a = [], []
    for i, p in enumerate(parts):
        ni, mi = p.n, p.m
        if not mi:
            continue
        w1 = gg.DeviceArray(ctx, mi, np.float32).fill(1.0)
        p.set_weights_dev(w1)
        w1.free()
        vi = gg.DeviceArray.from_host(ctx, v[vb[i]:vb[i + 1]])
        ui = gg.DeviceArray.from_host(ctx, u[vb[i]:vb[i + 1]])
        ri = gg.DeviceArray(ctx, ni, np.float32)
        gwi = gg.DeviceArray(ctx, mi, np.float32)
        gvi = gg.DeviceArray(ctx, ni, np.float32)
        gg.wpagerank(p, np.float32, 0.85, -1.0, iters, v=vi, out=ri)
        gg.wpagerank_adjoint(p, ri, ui, 0.85, -1.0, iters, gw=gwi, gv=gvi)
        t1 = time.perf_counter()
        gg.wpagerank(p, np.float32, 0.85, -1.0, iters, v=vi, out=ri)
        t2 = time.perf_counter()
        gg.wpagerank_adjoint(p, ri, ui, 0.85, -1.0, iters, gw=gwi, gv=gvi)
        t3 = time.perf_counter()
        per_f.append((t2 - t1) * 1e3)
        per_a.append((t3 - t2) * 1e3)
        for a in (vi, ui, ri, gwi, gvi):
            a.free()
    # the same requests as sub-batches of 64 graphs, one arena solve each
    # (forward + adjoint, timed after its capture)
    sub = 64
    sub_ms, sub_sizes = [], []
    for k in range(0, count, sub):
        sp = parts[k:k + sub]
        sa, _, _ = gg.Graph.arena(ctx, sp)
        lo, hi = int(vb[k]), int(vb[min(k + sub, count)])
        vs = gg.DeviceArray.from_host(ctx, v[lo:hi])
        us = gg.DeviceArray.from_host(ctx, u[lo:hi])
        rs = gg.DeviceArray(ctx, hi - lo, np.float32)
        gws = gg.DeviceArray(ctx, max(sa.m, 1), np.float32)
        gvs = gg.DeviceArray(ctx, hi - lo, np.float32)
        gg.wpagerank(sa, np.float32, 0.85, -1.0, iters, v=vs, out=rs)  # capture
        gg.wpagerank_adjoint(sa, rs, us, 0.85, -1.0, iters, gw=gws, gv=gvs)  # capture
        t1 = time.perf_counter()
        gg.wpagerank(sa, np.float32, 0.85, -1.0, iters, v=vs, out=rs)
        gg.wpagerank_adjoint(sa, rs, us, 0.85, -1.0, iters, gw=gws, gv=gvs)
        sub_ms.append((time.perf_counter() - t1) * 1e3)
        sub_sizes.append(len(sp))
        for a in (vs, us, rs, gws, gvs):
            a.free()
        sa.close()
    # Serving model: all 'count' requests arrive at t = 0 and are served
    # back to back; a request's latency is when its results are complete.
    # One batch: every graph completes with the single solve.  Sub-batches:
    # when its sub-batch's solve ends.  Graph by graph: when its own ends.
    T = tb_f + tb_a
    done_sub = np.repeat(np.cumsum(sub_ms), sub_sizes)
    done_seq = np.cumsum(np.add(per_f, per_a))
    M = arena.m
    res = {"graphs": count, "n": n, "m": M, "iterations": iters,
           "tuned": {"schedule": tune["schedule"], "edges_per_lane": tune["edges_per_lane"]},
           "batched_fwd_ms": tb_f, "batched_adj_ms": tb_a,
           "batched_fwd_gteps": iters * M / tb_f / 1e6, "batched_adj_gteps": iters * M / tb_a / 1e6,
           "batched_latency_ms": T,
           "unbatched_total_ms": float(np.sum(per_f) + np.sum(per_a)),
           "unbatched_fwd_gteps": iters * M / float(np.sum(per_f)) / 1e6,
           "unbatched_solve_ms_p50": float(np.percentile(np.add(per_f, per_a), 50)),
           "unbatched_solve_ms_p99": float(np.percentile(np.add(per_f, per_a), 99)),
           "serving": {
               "model": f"{count} requests at t=0, served back to back; latency = completion of the request's results",
               "one_batch": {"p50_ms": T, "p99_ms": T, "total_ms": T},
               f"sub_batches_of_{sub}": {"p50_ms": float(np.percentile(done_sub, 50)),
                                        "p99_ms": float(np.percentile(done_sub, 99)),
                                        "total_ms": float(np.sum(sub_ms)),
                                        "solve_ms": [round(x, 3) for x in sub_ms]},
               "graph_by_graph": {"p50_ms": float(np.percentile(done_seq, 50)),
                                  "p99_ms": float(np.percentile(done_seq, 99)),
                                  "total_ms": float(done_seq[-1])}}}
    arena.close()
    for p in parts:
        p.close()
    res.update(multisource_bench(gg, ctx, msbfs_scale, seed, plan))
    return res


def multisource_bench(gg, ctx, scale=22, seed=5, plan=True, calls=5):
    """32-source BFS on symmetric R-MAT 'scale' (edge factor 16) against 32
    single-source BFS runs, and BC.C's forward phase (levels and NumPaths,
    FP32) for the same 32 sources in one pass beside 32 single-source BC.C
    runs.  The multi-source times are the mean of 'calls' calls after one
    warm-up."""
    res = {}
    g = gg.Graph.rmat(ctx, scale, 16, seed, symmetric=True)
    g.partition(384)
    if plan:
        g.reorder(1)  # hot-sources-first numbering for both legs (untimed preprocessing)
    deg = g.degrees(False)
    src = np.random.default_rng(seed).choice(np.flatnonzero(deg > 0), 32, replace=False)
    lv, _ = gg.msbfs(g, src)  # warm-up (row blocks)
    t1 = time.perf_counter()
    for _ in range(calls):
        lv, info = gg.msbfs(g, src, out=lv)
    tm = (time.perf_counter() - t1) / calls
    L = lv.to_host().reshape(32, g.n)
    und = sum(int(deg[L[k] >= 0].astype(np.int64).sum()) // 2 for k in range(32))
    par = gg.DeviceArray(ctx, g.n, np.int32)
    gg.bfs(g, int(src[0]), parents=par, levels=False)  # warm-up (the plan's out-edge side)
    t1 = time.perf_counter()
    for s in src:
        gg.bfs(g, int(s), parents=par, levels=False)
    ts = time.perf_counter() - t1
    res["msbfs"] = {"scale": scale, "sources": 32, "levels": info.iters, "ms": tm * 1e3,
                    "gteps": und / tm / 1e9, "single_source_loop_ms": ts * 1e3, "single_source_gteps": und / ts / 1e9}
    # BC.C's forward phase (levels and NumPaths) for the same 32 sources in
    # one pass (FP32, the reference's fType), beside 32 single-source BC.C
    # runs (forward and backward) on the same graph
    pa, _, _ = gg.msbc_forward(g, src, np.float32)  # warm-up
    t1 = time.perf_counter()
    for _ in range(calls):
        pa, _, binfo = gg.msbc_forward(g, src, np.float32, paths=pa)
    tb = (time.perf_counter() - t1) / calls
    sc = gg.DeviceArray(ctx, g.n, np.float32)
    gg.bc(g, int(src[0]), np.float32, out=sc)  # warm-up
    t1 = time.perf_counter()
    for s in src:
        gg.bc(g, int(s), np.float32, out=sc)
    tbs = time.perf_counter() - t1
    res["msbc_forward"] = {"scale": scale, "sources": 32, "levels": binfo.iters, "ms": tb * 1e3,
                           "gteps": und / tb / 1e9, "single_source_bc_loop_ms": tbs * 1e3}
    g.close()
    return res


def apps_bench(gg, ctx, scale, steps, plan=True):
    """The further apps on the same engine: BC.C from root 100 (FP32, the
    reference's fType) on the symmetric R-MAT graph of the BFS leg, and
    BellmanFord.C from vertex 0 on a directed R-MAT graph (seed 4) with
    integer lengths 1..19; both on the hot-sources-first plan (untimed
    preprocessing) unless plan is False."""
    g = gg.Graph.rmat(ctx, scale, 16, 2, symmetric=True)
    if plan:
        g.reorder(1)
    sc = gg.DeviceArray(ctx, g.n, np.float32)
    gg.bc(g, 100, np.float32, out=sc)
    ts = []
    for _ in range(max(steps, 3)):
        _, _, info = gg.bc(g, 100, np.float32, out=sc)
        ts.append(info.ms)
    res = {"bc": {"scale": scale, "root": 100, "rounds": info.iters, "ms": float(np.median(ts)),
                  "paths": info.path_list(), "edges_read": int(info.edges_scanned),
                  "gbs": 8 * info.edges_scanned / float(np.median(ts)) / 1e6}}
    g.close()
    g = gg.Graph.rmat(ctx, scale - 2, 16, 4)
    w = (1 + (np.arange(g.m, dtype=np.uint64) * np.uint64(2654435761)) % np.uint64(19)).astype(np.float32)
    g.set_weights(w)
    if plan:
        g.reorder(1)
    d = gg.DeviceArray(ctx, g.n, np.int32)
    gg.sssp(g, 0, out=d)
    ts = []
    for _ in range(max(steps, 3)):
        _, info = gg.sssp(g, 0, out=d)
        ts.append(info.ms)
    res["sssp"] = {"scale": scale - 2, "rounds": info.iters, "ms": float(np.median(ts)),
                   "edges_read": int(info.edges_scanned)}
    g.close()
    return res


def dropin_bench(scale, rounds=3):
    """The reference's own app sources compiled unchanged against the drop-in
    front end (apps/_build/<App>), timed by their own driver ('Running : t'
    per round, ligra-numa.h:1155-1157), with the front end's per-round
    HBM<->host copy report (GG_FRONT_STATS): PageRank.C on config 1 (directed
    R-MAT 20) and on the headline graph, BFS.C and Components.C on symmetric
    R-MAT 'scale' (root 100), SPMV.C on directed R-MAT 'scale' (seed 3)."""
    runs = [("PageRank", ["-c", "384", "-rmat", "20", "16", "1"], "config 1, R-MAT 20"),
            ("PageRank", ["-c", "384", "-rmat", str(scale), "16", "1"], f"R-MAT {scale}"),
            ("BFS", ["-s", "-c", "384", "-r", "100", "-rmat", str(scale), "16", "2"], f"symmetric R-MAT {scale}"),
            ("Components", ["-s", "-c", "384", "-rmat", str(scale), "16", "2"], f"symmetric R-MAT {scale}"),
            ("SPMV", ["-c", "384", "-rmat", str(scale), "16", "3"], f"R-MAT {scale}")]
    out = []
    for app, args, what in runs:
        exe = os.path.join(ROOT, "apps", "_build", app)
        if not os.path.exists(exe):
            continue
        env = dict(os.environ, GG_FRONT_STATS="1")
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-rounds", str(rounds)] + args, capture_output=True, text=True, env=env,
                           timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            log(f"[bench] drop-in {app} failed: {r.stderr[-400:]}")
            continue
        times = [float(l.split(":")[1]) * 1e3 for l in r.stdout.splitlines() if l.startswith("Running")]
        stats = [l for l in r.stdout.splitlines() if l.startswith("gg_front: round")]
        rec = {"app": f"app/{app}.C", "graph": what, "round_ms": times,
               "median_round_ms": float(np.median(times)) if times else None, "wall_s": wall}
        if stats:
            # the last round's coherence traffic: copies each way, host loops
            last = stats[-1].split(":", 2)[2]
            rec["last_round_host_traffic"] = last.strip()
        out.append(rec)
        log(f"[bench] drop-in {app} ({what}): {rec}")
    return out


def host_cpu_record():
    """What the CPU baseline ran on: logical CPUs, the CPUs this process may
    use, the OpenMP thread setting and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "model": model}


def cpu_baseline(scale=24, rounds=3):
    """The reference engine (oracle/_ref, compiled from the reference's own
    sources with its OpenMP branch) on the headline workload: PageRank.C on
    directed R-MAT scale 24 (ef 16, seed 1), 384 partitions, 10 iterations,
    timed as the reference times Compute() (ligra-numa.h:1155-1157), on the
    host threads this process is given (OMP_NUM_THREADS where set, else every
    CPU in its affinity mask)."""
    import oracle

    hw = host_cpu_record()
    exe = os.path.join(oracle.REF_DIR, "ref_pagerank")
    threads = os.environ.get("OMP_NUM_THREADS") or str(hw["nproc"] or 1)
    env = dict(os.environ, OMP_NUM_THREADS=threads)
    if os.path.exists(exe):
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-rmat", str(scale), "16", "1", "-c", "384", "-rounds", str(rounds)],
                           capture_output=True, text=True, env=env, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode == 0:
            line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
            d = json.loads(line)
            t = float(np.mean(d["times"]))
            return {"value": 10 * d["m"] / t / 1e9, "unit": "GTEPS", "cores": int(d["threads"]),
                    "kind": "reference", "host": hw,
                    "round_ms": [x * 1e3 for x in d["times"]], "median_round_ms": float(np.median(d["times"])) * 1e3,
                    "sample": f"reference PageRank.C (OpenMP build of its own engine, NUMA off, PART96, Hilbert COO) "
                              f"on directed R-MAT scale {scale} ef 16 seed 1, n={d['n']} m={d['m']}, "
                              f"384 partitions, 10 iterations, mean of {rounds} rounds ({t * 1e3:.1f} ms/round), "
                              f"{d['threads']} OpenMP threads; {wall:.0f} s including graph build and partitioning"}
        log(f"[bench] reference harness failed: {r.stderr[-500:]}")
    # fallback: the single-threaded C restatement on a smaller graph
    off, ids = oracle.rmat_csr(18, 16, 1)
    og = oracle.Graph(off, ids, P=384)
    t1 = time.perf_counter()
    og.pagerank(10)
    t = time.perf_counter() - t1
    return {"value": 10 * len(ids) / t / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port", "host": hw,
            "sample": f"C restatement, R-MAT scale 18, 10 iterations ({t * 1e3:.0f} ms)"}


def copy_calibration(gg, ctx, nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size, for scaling the HBM counters
    (FETCH_SIZE / WRITE_SIZE per dispatch) of a profiled run to bytes."""
    a = gg.DeviceArray(ctx, nbytes // 4, np.uint32).fill(1)
    b = gg.DeviceArray(ctx, nbytes // 4, np.uint32)
    for _ in range(reps):
        gg.lib().gg_memcpy_d2d(ctx.h, b.ptr, a.ptr, nbytes)
    ctx.sync()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def traffic_from_profile(kernel, n, m):
    """HBM bytes per launch (per iteration for a multi-kernel iteration) of
    'kernel' on this workload, as recorded from the rocprofv3 counter runs
    in the machine-readable block of a profiles/*.md summary (None when no
    summary covers it)."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        rec = {}
        for line in open(os.path.join(pdir, name)):
            if line.startswith("    ") and "=" in line and not line.strip().startswith(("rocprofv3", "export")):
                k, v = line.strip().split("=", 1)
                rec[k.strip()] = v.strip()
        try:
            if rec.get("kernel") == kernel and int(rec["n"]) == n and int(rec["m"]) == m:
                return float(rec["traffic_bytes_per_launch"])
        except (KeyError, ValueError):
            continue
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the reference app sources run through the drop-in front end")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device copies (counter calibration under rocprofv3)")
    ap.add_argument("--only-bfs-cc", action="store_true",
                    help="only run the BFS + Components leg (for kernel traces of it)")
    ap.add_argument("--only-adjoints", action="store_true",
                    help="only run the SpMV and PageRank adjoint legs (for kernel traces of them)")
    ap.add_argument("--only-msbfs", action="store_true",
                    help="only run the 32-source BFS and BC forward leg (for kernel traces of it)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (--gpus 1): the engine runs on a single MI355X; "
              "sharding across GPUs is designed but not built or measured.", file=sys.stderr)
        sys.exit(2)
    import graphgrind_amd as gg

    ctx = gg.Context(0)
    if a.calibrate_copy:
        copy_calibration(gg, ctx)
        ctx.close()
        return
    if a.only_bfs_cc:
        log(f"[bench] bfs/cc: {bfs_cc_bench(gg, ctx, a.scale, a.steps)}")
        ctx.close()
        return
    if a.only_msbfs:
        log(f"[bench] multi-source: {multisource_bench(gg, ctx, a.scale - 2)}")
        ctx.close()
        return
    if a.only_adjoints:
        log(f"[bench] spmv adjoint: {spmv_adjoint_bench(gg, ctx, a.scale, a.steps)}")
        log(f"[bench] pagerank adjoint: {wpagerank_adjoint_bench(gg, ctx, a.scale - 2, a.steps)}")
        ctx.close()
        return
    pr = pagerank_bench(gg, ctx, a.scale, a.steps, a.warmup, np.float64)
    log(f"[bench] pagerank fp64: {pr}")
    extras = {}
    if not a.no_extras:
        pr32 = pagerank_bench(gg, ctx, a.scale, a.steps, a.warmup, np.float32)
        log(f"[bench] pagerank fp32: {pr32}")
        extras["pagerank_fp32"] = {k: pr32[k] for k in ("gteps", "step_ms", "kernel_ms", "achieved_gbs", "kernel")}
        bc = bfs_cc_bench(gg, ctx, a.scale, a.steps)
        log(f"[bench] bfs/cc: {bc}")
        extras["bfs_cc_sym_rmat"] = bc
        sa = spmv_adjoint_bench(gg, ctx, a.scale, a.steps)
        log(f"[bench] spmv adjoint: {sa}")
        extras["spmv_adjoint_fp32"] = sa
        wa = wpagerank_adjoint_bench(gg, ctx, a.scale - 2, a.steps)
        log(f"[bench] pagerank adjoint: {wa}")
        extras["pagerank_adjoint_fp64"] = wa
        ap_ = apps_bench(gg, ctx, a.scale, a.steps)
        log(f"[bench] apps: {ap_}")
        extras["apps"] = ap_
        if not a.no_batch:
            bb = batched_bench(gg, ctx, msbfs_scale=a.scale - 2)
            log(f"[bench] batched: {bb}")
            extras["batched"] = bb
    ctx.close()
    if not a.no_extras and not a.no_dropin:
        extras["dropin_apps"] = dropin_bench(a.scale)
    cpu = None if a.no_cpu else cpu_baseline(a.scale)
    line = {
        "metric": "GTEPS & % HBM peak (PageRank/BFS, RMAT-24); adjoint/forward time ratio",
        "value": pr["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": pr["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded R-MAT generated on device)",
        "config": {"workload": f"PageRank.C (frontier-gated, 10 iterations) on directed R-MAT scale {a.scale}, "
                               f"edge factor 16, seed 1, 384 edge-balanced Hilbert COO partitions, FP64",
                   "n": pr["n"], "m": pr["m"], "iterations": 10, "partitions": 384},
        "roofline": {"bound": "hbm", "achieved": pr["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": pr["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(pr["kernel"], pr["n"], pr["m"]),
                     "kernel": pr["kernel"], "kernel_ms": pr["kernel_ms"],
                     "algo_bytes_per_launch": pr["algo_bytes_per_launch"]},
        "timing": {"mean_step_ms": pr["step_ms"], "median_compute_event_ms": pr["median_event_ms"],
                   "first_call_ms_incl_allocation": pr["first_call_ms"],
                   "note": "a step is one Compute(): initialisation + 10 iterations, buffers kept between "
                           "calls; the first call on the graph also allocates them and builds the "
                           "source-segmented layout of the plan's in-edges"},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
