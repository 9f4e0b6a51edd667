"""GPU parity of the differentiated forms against the FP64 oracle
(oracle/gg_oracle_grad.c), plus size-independent identities at full size
and the torch.autograd gradient-check harness.

Tolerances: rel L1 <= 1e-12 in FP64 (same iteration counts as the oracle),
<= 1e-5 in FP32 for fixed iteration counts; gradients within 1e-6 relative
of the FP64 oracle.
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
ALPHA = 0.85


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _weighted(ctx, scale, seed, ef=16):
    g = gg.Graph.rmat(ctx, scale, ef, seed, weighted=True)
    off, ids, w = g.csr()
    return g, off, ids, w


# ---------------------------------------------------------------- SpMV --

@pytest.mark.parametrize("scale", [10, 16])
def test_spmv_forward_weighted_vs_oracle(ctx, scale):
    print(f"spmv fwd scale={scale}", flush=True)
    g, off, ids, w = _weighted(ctx, scale, 3)
    x = oracle.uniform_pm1(3, g.n)
    want = oracle.spmv_csr(off, ids, w, x)
    for dt, tol in ((np.float64, TOL64), (np.float32, TOL32)):
        xd = gg.DeviceArray.from_host(ctx, x.astype(dt))
        y, _ = gg.spmv(g, xd)
        assert _rel(y.to_host(), want) < tol, dt
    g.close()


@pytest.mark.parametrize("scale", [1, 10, 16])
def test_spmv_bwd_vs_oracle(ctx, scale):
    print(f"spmv bwd scale={scale}", flush=True)
    g, off, ids, w = _weighted(ctx, scale, 3)
    n = g.n
    x = oracle.uniform_pm1(3, n)
    t = np.random.default_rng(scale).standard_normal(n)
    y = oracle.spmv_csr(off, ids, w, x)
    gy = y - t
    gx_w, gw_w = oracle.spmv_bwd(off, ids, w, x, gy)
    for dt, tol in ((np.float64, TOL64), (np.float32, TOL32)):
        xd = gg.DeviceArray.from_host(ctx, x.astype(dt))
        gyd = gg.DeviceArray.from_host(ctx, gy.astype(dt))
        gx, gw, _ = gg.spmv_bwd(g, xd, gyd)
        assert _rel(gx.to_host(), gx_w) < tol, dt
        if g.m:
            gwh = gw.to_host()[: g.m]
            if dt == np.float64:
                assert np.array_equal(gwh, gw_w)  # one product per edge: exact
            else:
                assert _rel(gwh, gw_w) < tol
        # gx alone and gw alone give the same values
        gx2, _, _ = gg.spmv_bwd(g, None, gyd, want_gw=False)
        assert np.array_equal(gx2.to_host(), gx.to_host())
    g.close()


def test_spmv_bwd_ragged_rows(ctx):
    """Rows far longer than a tile, empty rows and a tail of isolated
    vertices: the edge pass finds every edge's row across tile cuts."""
    n = 5000
    src = np.concatenate([np.zeros(4000, np.int64), np.full(3000, 7), np.arange(10, 2000)])
    dst = np.concatenate([np.arange(1, 4001), np.arange(100, 3100), np.arange(11, 2001) % n])
    off, ids = oracle.csr_from_edges(n, src, dst)
    w = oracle.weights(9, len(ids))
    g = gg.Graph.from_csr(ctx, off, ids, w=w)
    x = oracle.uniform_pm1(5, n)
    gy = oracle.uniform_pm1(6, n)
    gx_w, gw_w = oracle.spmv_bwd(off, ids, w, x, gy)
    gx, gw, _ = gg.spmv_bwd(g, gg.DeviceArray.from_host(ctx, x), gg.DeviceArray.from_host(ctx, gy))
    assert _rel(gx.to_host(), gx_w) < TOL64
    assert np.array_equal(gw.to_host()[: g.m], gw_w)
    g.close()


@pytest.mark.parametrize("scale", [12, 17])
@pytest.mark.parametrize("schedule", [1, 0])
def test_spmv_renumbered_plan(ctx, scale, schedule):
    """With the hot-sources-first plan, the forward pulls over the relabelled
    in-edge twin and the backward over the relabelled CSR (x, gy permuted in,
    y, gx scattered back, dL/dw written at the caller's edge positions).
    Only the summation order changes: oracle bars as without the plan, dL/dw
    exact in FP64, bitwise run to run, gx alone == gx of the fused pass, and
    dropping the plan restores the plain results bit for bit."""
    print(f"spmv plan scale={scale} schedule={schedule}", flush=True)
    g, off, ids, w = _weighted(ctx, scale, 3)
    g.set_option(gg.OPT_PULL_SCHEDULE, schedule)
    n = g.n
    x = oracle.uniform_pm1(3, n)
    gy = np.random.default_rng(scale).standard_normal(n)
    want_y = oracle.spmv_csr(off, ids, w, x)
    want_gx, want_gw = oracle.spmv_bwd(off, ids, w, x, gy)
    for dt, tol in ((np.float64, TOL64), (np.float32, TOL32)):
        xd = gg.DeviceArray.from_host(ctx, x.astype(dt))
        gyd = gg.DeviceArray.from_host(ctx, gy.astype(dt))
        outs = []
        for mode in (0, 1, 1, 0):
            g.reorder(mode)
            y, _ = gg.spmv(g, xd)
            gx, gw, _ = gg.spmv_bwd(g, xd, gyd)
            gx2, _, _ = gg.spmv_bwd(g, None, gyd, want_gw=False)
            outs.append([a.to_host() for a in (y, gx, gw, gx2)])
        for a, b in zip(outs[0], outs[3]):
            assert np.array_equal(a, b), dt
        for a, b in zip(outs[1], outs[2]):
            assert np.array_equal(a, b), dt
        y1, gx1, gw1, gx21 = outs[1]
        assert np.array_equal(gx1, gx21)
        assert _rel(y1, want_y) < tol and _rel(gx1, want_gx) < tol
        if dt == np.float64:
            assert np.array_equal(gw1[: g.m], want_gw)
        else:
            assert np.array_equal(gw1[: g.m], outs[0][2][: g.m])  # same FP32 products
    g.close()


def test_spmv_plan_ragged_rows(ctx):
    """The relabelled CSR's edge-position map across empty rows, a row far
    longer than a work item and a tail of isolated vertices."""
    n = 5000
    src = np.concatenate([np.zeros(4000, np.int64), np.full(3000, 7), np.arange(10, 2000)])
    dst = np.concatenate([np.arange(1, 4001), np.arange(100, 3100), np.arange(11, 2001) % n])
    off, ids = oracle.csr_from_edges(n, src, dst)
    w = oracle.weights(9, len(ids))
    g = gg.Graph.from_csr(ctx, off, ids, w=w).reorder(1)
    x = oracle.uniform_pm1(5, n)
    gy = oracle.uniform_pm1(6, n)
    gx_w, gw_w = oracle.spmv_bwd(off, ids, w, x, gy)
    xd = gg.DeviceArray.from_host(ctx, x)
    gx, gw, _ = gg.spmv_bwd(g, xd, gg.DeviceArray.from_host(ctx, gy))
    assert _rel(gx.to_host(), gx_w) < TOL64
    assert np.array_equal(gw.to_host()[: g.m], gw_w)
    y, _ = gg.spmv(g, xd)
    assert _rel(y.to_host(), oracle.spmv_csr(off, ids, w, x)) < TOL64
    g.close()


def test_plan_follows_new_weights(ctx):
    """Weights set after the plan was built (on an unweighted graph, then
    changed again) reach the plan's copies: SpMV and weighted PageRank
    with the plan match the oracle on the new weights."""
    off, ids = oracle.rmat_csr(11, 8, 12)
    n, m = len(off) - 1, len(ids)
    g = gg.Graph.from_csr(ctx, off, ids).reorder(1)
    x = oracle.uniform_pm1(4, n)
    xd = gg.DeviceArray.from_host(ctx, x)
    for seed in (21, 22):
        w = oracle.weights(seed, m)
        g.set_weights(w)
        y, _ = gg.spmv(g, xd)
        assert _rel(y.to_host(), oracle.spmv_csr(off, ids, w, x)) < TOL64, seed
        want, it, _ = oracle.wpagerank(off, ids, w, ALPHA, 1e-7, 100)
        r, info = gg.wpagerank(g, np.float64, ALPHA, 1e-7, 100)
        assert info.iters == it and _rel(r.to_host(), want) < TOL64, seed
    g.close()


def test_spmv_bwd_euler_identities_scale24(ctx):
    """At the full SpMV size (R-MAT scale 24, seed 3, FP32): y is linear in
    x and in w, so x.gx = w.gw = gy.y for any gy."""
    print("spmv identities scale=24", flush=True)
    g = gg.Graph.rmat(ctx, 24, 16, 3, weighted=True)
    n, m = g.n, g.m
    x = gg.DeviceArray.from_host(ctx, oracle.uniform_pm1(3, n).astype(np.float32))
    y, _ = gg.spmv(g, x)
    gy = gg.DeviceArray.from_host(ctx, np.random.default_rng(3).standard_normal(n).astype(np.float32))
    gx, gw, _ = gg.spmv_bwd(g, x, gy)
    xh, yh, gyh, gxh = (a.to_host().astype(np.float64) for a in (x, y, gy, gx))
    _, _, w = g.csr()
    gwh = gw.to_host()[:m].astype(np.float64)
    ref = float(np.dot(gyh, yh))
    assert abs(np.dot(xh, gxh) - ref) <= 1e-4 * np.abs(gyh * yh).sum()
    assert abs(np.dot(w.astype(np.float64), gwh) - ref) <= 1e-4 * np.abs(gyh * yh).sum()
    g.close()


# --------------------------------------------------------- weighted PR --

@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale", [8, 14, 18])
def test_wpagerank_fp64_vs_oracle(ctx, scale, plan):
    """plan 1: the solve runs in the hot-sources-first numbering (relabelled
    twin, vectors permuted in and out); only the summation order changes."""
    print(f"wpagerank fp64 scale={scale} plan={plan}", flush=True)
    g, off, ids, w = _weighted(ctx, scale, 4)
    g.reorder(plan)
    want, it, _ = oracle.wpagerank(off, ids, w, ALPHA, 1e-7, 100)
    r, info = gg.wpagerank(g, np.float64, ALPHA, 1e-7, 100)
    assert info.iters == it
    assert _rel(r.to_host(), want) < TOL64
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_wpagerank_personalised_and_fixed_iterations(ctx, plan):
    g, off, ids, w = _weighted(ctx, 12, 4)
    g.reorder(plan)
    n = g.n
    v = np.zeros(n)
    v[[3, 99, 1234, 4000]] = 0.25
    want, it, _ = oracle.wpagerank(off, ids, w, ALPHA, 1e-9, 100, v=v)
    r, info = gg.wpagerank(g, np.float64, ALPHA, 1e-9, 100, v=gg.DeviceArray.from_host(ctx, v))
    assert info.iters == it and _rel(r.to_host(), want) < TOL64
    # FP32, exactly 30 iterations (eps 0), against the FP64 oracle
    want30, it30, _ = oracle.wpagerank(off, ids, w, ALPHA, 0.0, 30)
    r32, info32 = gg.wpagerank(g, np.float32, ALPHA, 0.0, 30)
    assert info32.iters == 30 == it30
    assert _rel(r32.to_host(), want30) < TOL32
    g.close()


def _targets(n, k, seed):
    return np.random.default_rng(seed).choice(n, min(k, n), replace=False)


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale", [8, 14, 18])
def test_wpagerank_adjoint_fp64_vs_oracle(ctx, scale, plan):
    """Gradients w.r.t. the edge weights and the teleport vector, at the
    GPU's own forward output, against the FP64 oracle (plan 1: both solves
    in the hot-sources-first numbering, over the relabelled twin and CSR)."""
    print(f"wpagerank adjoint fp64 scale={scale} plan={plan}", flush=True)
    g, off, ids, w = _weighted(ctx, scale, 4)
    g.reorder(plan)
    n = g.n
    r, _ = gg.wpagerank(g, np.float64, ALPHA, 1e-7, 100)
    rh = r.to_host()
    T = _targets(n, 4096, scale)
    u = np.zeros(n)
    u[T] = 2 * rh[T]
    lam_w, gw_w, gv_w, it = oracle.wpagerank_adjoint(off, ids, w, rh, u, ALPHA, 1e-7, 100)
    gw, gv, lam, info = gg.wpagerank_adjoint(g, r, gg.DeviceArray.from_host(ctx, u), ALPHA, 1e-7, 100,
                                             want_lam=True)
    assert info.iters == it
    assert _rel(lam.to_host(), lam_w) < TOL64
    assert _rel(gv.to_host(), gv_w) < TOL64
    assert _rel(gw.to_host()[: g.m], gw_w) < TOL64
    # per-coordinate on the large entries: within 1e-6 relative
    gwh = gw.to_host()[: g.m]
    big = np.abs(gw_w) > 1e-3 * np.abs(gw_w).max()
    assert np.all(np.abs(gwh[big] - gw_w[big]) <= 1e-6 * np.abs(gw_w[big]))
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_wpagerank_adjoint_identities_scale22(ctx, plan):
    """The differentiable-PageRank shape (R-MAT scale 22, ef 16, seed 4,
    4096 targets, FP64): Euler's identities hold at any size --
    v.dL/dv = 2L (r is linear in v) and, per source, sum_e w_e dL/dw_e = 0
    (P does not change when one source's out-weights are scaled)."""
    print(f"wpagerank identities scale=22 plan={plan}", flush=True)
    g = gg.Graph.rmat(ctx, 22, 16, 4, weighted=True)
    g.reorder(plan)
    n, m = g.n, g.m
    r, finfo = gg.wpagerank(g, np.float64, ALPHA, 1e-7, 100)
    rh = r.to_host()
    T = _targets(n, 4096, 22)
    u = np.zeros(n)
    u[T] = 2 * rh[T]
    L = float(np.sum(rh[T] ** 2))
    gw, gv, _, ainfo = gg.wpagerank_adjoint(g, r, gg.DeviceArray.from_host(ctx, u), ALPHA, 1e-7, 100)
    assert 1 < finfo.iters < 100 and 1 < ainfo.iters < 100
    gvh = gv.to_host()
    # both solves stop at 1e-7 steps, which bounds the identity's defect
    assert abs(np.dot(np.full(n, 1.0 / n), gvh) - 2 * L) <= 1e-5 * 2 * L
    off, ids, w = g.csr()
    gwh = gw.to_host()[:m]
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    per_src = np.bincount(src, weights=w.astype(np.float64) * gwh, minlength=n)
    scale = np.bincount(src, weights=np.abs(w.astype(np.float64) * gwh), minlength=n)
    # exact at the fixed point; the stopped adjoint leaves a defect of the
    # order of its last step (largest, relatively, for one-edge sources)
    assert np.abs(per_src).sum() <= 1e-5 * scale.sum()
    g.close()


@pytest.mark.parametrize("dt,tol", [(np.float64, TOL64), (np.float32, TOL32)])
def test_wpagerank_fixed_loops_with_plan(ctx, dt, tol):
    """The captured fixed-iteration loops (eps < 0) in the plan's numbering:
    against the oracle and the unplanned loops, replayed twice (bitwise the
    same), and re-captured after the plan is dropped."""
    g, off, ids, w = _weighted(ctx, 13, 4)
    n = g.n
    v = np.zeros(n)
    v[[5, 77, 2048]] = 1 / 3
    want, it, _ = oracle.wpagerank(off, ids, w, ALPHA, 0.0, 25, v=v)
    vd = gg.DeviceArray.from_host(ctx, v.astype(dt))
    u = np.zeros(n)
    T = _targets(n, 64, 13)
    u[T] = 1.0
    ud = gg.DeviceArray.from_host(ctx, u.astype(dt))
    res = {}
    for plan in (1, 0):
        g.reorder(plan)
        outs = []
        for _ in range(2):
            r, info = gg.wpagerank(g, dt, ALPHA, -1.0, 25, v=vd)
            assert info.iters == 25 == it
            gw, gv, lam, ainfo = gg.wpagerank_adjoint(g, r, ud, ALPHA, -1.0, 25, want_lam=True)
            assert ainfo.iters == 25
            outs.append([a.to_host() for a in (r, gw, gv, lam)])
        for a, b in zip(*outs):
            assert np.array_equal(a, b)
        assert _rel(outs[0][0], want) < tol
        res[plan] = outs[0]
    lam_w, gw_w, gv_w, _ = oracle.wpagerank_adjoint(off, ids, w, res[0][0].astype(np.float64), u, ALPHA, 0.0, 25)
    for plan in (0, 1):
        r_, gw_, gv_, lam_ = res[plan]
        assert _rel(r_, res[0][0]) < tol
        assert _rel(lam_, lam_w) < tol and _rel(gv_, gv_w) < tol
        if dt == np.float64:  # gw = coef (lam_d - mu_s) cancels: FP32 is not held to 1e-5
            assert _rel(gw_[: g.m], gw_w) < tol
    g.close()


def test_adjoint_errors(ctx):
    g = gg.Graph.rmat(ctx, 8, 16, 4)  # unweighted
    r = gg.DeviceArray(ctx, g.n, np.float64)
    with pytest.raises(gg.GGError) as e:
        gg.wpagerank(g, np.float64, out=r)
    assert e.value.code == -7
    with pytest.raises(gg.GGError):
        gg.wpagerank(g, np.float64, alpha=1.0, out=r)
    g.close()


# ------------------------------------------------ torch.autograd harness --

torch = pytest.importorskip("torch")


def test_autograd_spmv_gradcheck(ctx):
    from autograd_harness import SpmvFn

    g, off, ids, w = _weighted(ctx, 6, 3)
    x = torch.tensor(oracle.uniform_pm1(3, g.n), dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: SpmvFn.apply(t, g), (x,), eps=1e-6, atol=1e-9, rtol=1e-7)
    g.close()


def test_autograd_spmv_weight_grad_vs_fd(ctx):
    """dL/dw for L = 0.5|y - t|^2 through torch.autograd, against central
    differences with exactly representable FP32 steps (L is quadratic in
    w, so the difference quotient is exact up to rounding)."""
    from autograd_harness import SpmvWeightFn

    g, off, ids, w = _weighted(ctx, 8, 3)
    n, m = g.n, g.m
    x = torch.tensor(oracle.uniform_pm1(3, n), dtype=torch.float64, device="cuda")
    t = torch.tensor(np.random.default_rng(1).standard_normal(n), dtype=torch.float64, device="cuda")
    wt = torch.tensor(w, dtype=torch.float32, device="cuda", requires_grad=True)

    def loss(wv):
        y = SpmvWeightFn.apply(wv, x, g)
        return 0.5 * ((y - t) ** 2).sum()

    (gw,) = torch.autograd.grad(loss(wt), wt)
    gw = gw.cpu().numpy().astype(np.float64)
    for e in np.random.default_rng(2).choice(m, 16, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] = np.float32(w[e] + np.float32(2.0 ** -10))
        wm[e] = np.float32(w[e] - np.float32(2.0 ** -10))
        h = float(wp[e]) - float(wm[e])
        with torch.no_grad():
            lp = float(loss(torch.tensor(wp, device="cuda")))
            lm = float(loss(torch.tensor(wm, device="cuda")))
        fd = (lp - lm) / h
        assert abs(fd - gw[e]) <= 1e-5 * max(1.0, abs(fd)), (e, fd, gw[e])
    g.close()


def test_autograd_wpagerank_gradcheck_teleport(ctx):
    """torch.autograd.gradcheck of r(v) on a small graph, forward and
    adjoint solved to machine precision."""
    from autograd_harness import WPageRankFn

    g, off, ids, w = _weighted(ctx, 5, 4)
    wt = torch.tensor(w, dtype=torch.float32, device="cuda")
    v = torch.full((g.n,), 1.0 / g.n, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = lambda vv: WPageRankFn.apply(wt, vv, g, ALPHA, 0.0, 400, 0.0)  # noqa: E731
    assert torch.autograd.gradcheck(fn, (v,), eps=1e-6, atol=1e-8, rtol=1e-6)
    g.close()


def test_autograd_wpagerank_weight_grad_vs_fd(ctx):
    """dL/dw for L = sum_T r_t^2 through torch.autograd against central
    differences of the forward solve (representable FP32 steps) and the
    FP64 oracle gradient."""
    from autograd_harness import WPageRankFn

    g, off, ids, w = _weighted(ctx, 9, 4)
    n, m = g.n, g.m
    T = torch.tensor(_targets(n, 32, 9), device="cuda")
    v = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda")
    wt = torch.tensor(w, dtype=torch.float32, device="cuda", requires_grad=True)

    def loss(wv):
        r = WPageRankFn.apply(wv, v, g, ALPHA, 0.0, 400, 0.0)
        return (r[T] ** 2).sum()

    L = loss(wt)
    (gw,) = torch.autograd.grad(L, wt)
    gw = gw.cpu().numpy().astype(np.float64)
    rh, _, _ = oracle.wpagerank(off, ids, w, ALPHA, 0.0, 400)
    u = np.zeros(n)
    Th = T.cpu().numpy()
    u[Th] = 2 * rh[Th]
    _, gw_w, _, _ = oracle.wpagerank_adjoint(off, ids, w, rh, u, ALPHA, 0.0, 400)
    # gw went through FP32 on its way to w.grad
    assert _rel(gw, gw_w) < 1e-6
    gmax = np.abs(gw_w).max()
    for e in np.argsort(-np.abs(gw_w))[:12]:
        wp, wm = w.copy(), w.copy()
        wp[e] = np.float32(w[e] + np.float32(2.0 ** -12))
        wm[e] = np.float32(w[e] - np.float32(2.0 ** -12))
        h = float(wp[e]) - float(wm[e])
        with torch.no_grad():
            fd = (float(loss(torch.tensor(wp, device="cuda"))) - float(loss(torch.tensor(wm, device="cuda")))) / h
        assert abs(fd - gw_w[e]) <= 1e-4 * gmax, (e, fd, gw_w[e])
    g.set_weights(w)
    g.close()
