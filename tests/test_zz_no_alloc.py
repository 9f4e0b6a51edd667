"""gg_spmv_noalloc (the entry point that allocates nothing) against gg_spmv:
the same kernels, so the results must be bitwise equal.  Two calls in a row,
the second on a smaller graph (the shared scratch array keeps the first
call's contents), in FP32 and FP64, with and without a renumbered plan, and
on a plan that takes the source-segmented pull; and the documented refusals
(too large for the scratch array, schedule or segmented layout not built)
come back as status codes before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

import graphgrind_amd as gg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _noalloc(g, dt, x, y):
    L = gg.lib()
    L.gg_spmv_noalloc.restype = C.c_int
    L.gg_spmv_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    rc = L.gg_spmv_noalloc(g.h, gg.DT_OF[np.dtype(dt)], C.c_void_p(x.ptr), C.c_void_p(y.ptr))
    if rc == 0:
        g.ctx.sync()
    return rc


def _case(ctx, scale, seed, dt, plan):
    g = gg.Graph.rmat(ctx, scale, 16, seed, weighted=True)
    if plan:
        g.reorder(1)
    x = gg.DeviceArray.from_host(ctx, np.random.default_rng(seed).uniform(-1, 1, g.n).astype(dt))
    want, _ = gg.spmv(g, x)  # also builds the row-block schedule of an unplanned graph
    y = gg.DeviceArray(ctx, g.n, dt).fill(7.0)
    assert _noalloc(g, dt, x, y) == 0, gg.lib().gg_last_error().decode()
    got = y.to_host()
    assert np.array_equal(got, want.to_host()), (scale, np.dtype(dt).name, plan)
    g.close()


@pytest.mark.parametrize("plan", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_matches_spmv_two_calls(ctx, dt, plan):
    print(f"no-alloc spmv {np.dtype(dt).name} plan={plan}", flush=True)
    _case(ctx, 18, 3, dt, plan)  # first call
    _case(ctx, 14, 4, dt, plan)  # second call, smaller graph


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_segmented_pull(ctx, dt):
    """On a plan whose gg_spmv takes the source-segmented pull (forced here
    with GG_OPT_XCD_SEGMENTS = 2), the entry point takes it too: refused
    until one gg_spmv call has built the layout, then equal to gg_spmv bit
    for bit, on two graphs in a row, the second smaller."""
    for scale, seed in ((16, 5), (13, 6)):
        print(f"no-alloc segmented spmv {np.dtype(dt).name} scale={scale}", flush=True)
        g = gg.Graph.rmat(ctx, scale, 16, seed, weighted=True)
        g.reorder(1)
        g.set_option(3, 2)  # GG_OPT_XCD_SEGMENTS: always
        x = gg.DeviceArray.from_host(ctx, np.random.default_rng(seed).uniform(-1, 1, g.n).astype(dt))
        y = gg.DeviceArray(ctx, g.n, dt).fill(7.0)
        assert _noalloc(g, dt, x, y) == -7  # GG_ERR_STATE: no layout yet
        want, _ = gg.spmv(g, x)
        assert _noalloc(g, dt, x, y) == 0, gg.lib().gg_last_error().decode()
        assert np.array_equal(y.to_host(), want.to_host()), (scale, np.dtype(dt).name)
        g.close()


def test_noalloc_refusals(ctx):
    # FP64 x of R-MAT 24 (134 MB) does not fit the 64 MiB array
    g = gg.Graph.rmat(ctx, 24, 16, 3, weighted=True)
    g.reorder(1)
    x = gg.DeviceArray(ctx, g.n, np.float64).fill(1.0)
    y = gg.DeviceArray(ctx, g.n, np.float64)
    assert _noalloc(g, np.float64, x, y) == -4  # GG_ERR_SIZE
    g.close()
    # an unplanned graph whose row-block schedule no pull has built yet
    g = gg.Graph.rmat(ctx, 12, 16, 3, weighted=True)
    x = gg.DeviceArray(ctx, g.n, np.float32).fill(1.0)
    y = gg.DeviceArray(ctx, g.n, np.float32)
    assert _noalloc(g, np.float32, x, y) == -7  # GG_ERR_STATE
    g.close()
