"""The context's buffer pool (frontier arrays and per-call temporaries are
reused instead of freed): repeated traversals keep device memory flat, and a
frontier released after its context is gone is freed, not pooled."""
import ctypes as C

import numpy as np
import pytest

import graphgrind_amd as gg

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_pool_keeps_memory_flat():
    ctx = gg.Context(0)
    g = gg.Graph.rmat(ctx, 16, 16, 2, symmetric=True).partition(64)
    lab = gg.DeviceArray(ctx, g.n, np.uint32)

    def work():
        gg.bfs(g, 100)
        gg.components(g, out=lab)
        f = gg.Frontier.single(g, 5)
        out, _ = gg.edge_map(g, f, gg.OP_CC, lab, lab)
        out.close()
        f.close()

    for _ in range(3):
        work()
    ctx.sync()
    before = _free_bytes()
    for _ in range(20):
        work()
    ctx.sync()
    after = _free_bytes()
    assert after >= before - (16 << 20), (before, after)  # no growth beyond noise
    g.close()
    ctx.close()


def test_frontier_destroyed_after_its_context():
    ctx = gg.Context(0)
    g = gg.Graph.rmat(ctx, 10, 8, 1).partition(8)
    f = gg.Frontier.bits(g)
    out, _ = gg.edge_map(g, f, gg.OP_PAGERANK, gg.DeviceArray(ctx, g.n, np.float64).fill(1.0 / g.n),
                         gg.DeviceArray(ctx, g.n, np.float64).fill(0.0))
    ctx.sync()
    g.close()
    assert gg.lib().gg_ctx_destroy(ctx.h) == 0
    ctx.h = None
    # the frontiers' arrays are freed directly: their context is gone
    assert gg.lib().gg_frontier_destroy(out.h) == 0
    assert gg.lib().gg_frontier_destroy(f.h) == 0
    out.h = None
    f.h = None
