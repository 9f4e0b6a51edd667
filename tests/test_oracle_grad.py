"""Pins for the oracle's differentiated forms (oracle/gg_oracle_grad.c).

Each check is an independent formulation: scipy sparse products and direct
solves of the fixed-point equations, dense brute-force derivatives of the
transition matrix on tiny graphs, central finite differences of the loss,
and homogeneity identities (Euler's theorem) that a dropped term, a wrong
index or a transposed operand breaks.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle

ALPHA = 0.85


def _graph(scale, seed, ef=8):
    off, ids = oracle.rmat_csr(scale, ef, seed)
    w = oracle.weights(seed, len(ids))
    return off, ids, w


def _adj(off, ids, w):
    """A[s, d] = w_{s->d} as FP64 CSR."""
    n = len(off) - 1
    return sp.csr_matrix((np.asarray(w, np.float64), ids.astype(np.int64), off.astype(np.int64)), shape=(n, n))


def _trans(off, ids, w):
    """P[d, s] = w_{s->d} / S_s (dangling columns zero)."""
    A = _adj(off, ids, w)
    S = np.asarray(A.sum(axis=1)).ravel()
    inv = np.divide(1.0, S, out=np.zeros_like(S), where=S > 0)
    return (sp.diags(inv) @ A).T.tocsr(), S


def _solve_r(P, v):
    n = P.shape[0]
    return spla.spsolve((sp.identity(n, format="csc") - ALPHA * P).tocsc(), (1 - ALPHA) * v)


def _solve_lam(P, u):
    n = P.shape[0]
    return spla.spsolve((sp.identity(n, format="csc") - ALPHA * P.T).tocsc(), u)


def _sources(off):
    """Source vertex of every CSR edge."""
    return np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))


def _rel(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


# ---------------------------------------------------------------- SpMV --

@pytest.mark.parametrize("scale,seed", [(6, 3), (11, 3)])
def test_spmv_forward_and_adjoint_vs_scipy(scale, seed):
    off, ids, w = _graph(scale, seed)
    n = len(off) - 1
    x = oracle.uniform_pm1(seed, n)
    A = _adj(off, ids, w)
    y = oracle.spmv_csr(off, ids, w, x)
    assert _rel(y, A.T @ x) < 1e-13
    gy = np.random.default_rng(seed).standard_normal(n)
    gx, gw = oracle.spmv_bwd(off, ids, w, x, gy)
    assert _rel(gx, A @ gy) < 1e-13
    src = _sources(off)
    assert np.array_equal(gw, x[src] * gy[ids])
    # Euler: y is linear in w and in x
    assert np.isclose(np.dot(np.asarray(w, np.float64), gw), np.dot(gy, y), rtol=1e-12)
    assert np.isclose(np.dot(x, gx), np.dot(gy, y), rtol=1e-12)


def test_spmv_adjoint_central_differences():
    """L = 0.5 |y - t|^2 is quadratic in w and x: central differences with
    exactly representable FP32 steps equal the derivative up to rounding."""
    off, ids, w = _graph(7, 3)
    n, m = len(off) - 1, len(ids)
    x = oracle.uniform_pm1(3, n)
    t = np.random.default_rng(1).standard_normal(n)

    def loss(wv, xv):
        y = oracle.spmv_csr(off, ids, wv, xv)
        return 0.5 * float(np.sum((y - t) ** 2))

    y = oracle.spmv_csr(off, ids, w, x)
    gx, gw = oracle.spmv_bwd(off, ids, w, x, y - t)
    rng = np.random.default_rng(2)
    for e in rng.choice(m, 24, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] = np.float32(w[e] + np.float32(2.0 ** -10))
        wm[e] = np.float32(w[e] - np.float32(2.0 ** -10))
        h = float(wp[e]) - float(wm[e])
        fd = (loss(wp, x) - loss(wm, x)) / h
        assert abs(fd - gw[e]) <= 1e-7 * max(1.0, abs(gw[e])), (e, fd, gw[e])
    for s in rng.choice(n, 24, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[s] += 1e-3
        xm[s] -= 1e-3
        fd = (loss(w, xp) - loss(w, xm)) / 2e-3
        assert abs(fd - gx[s]) <= 1e-7 * max(1.0, abs(gx[s]))


# --------------------------------------------------------- weighted PR --

@pytest.mark.parametrize("scale,seed", [(8, 4), (12, 4)])
def test_wpagerank_fixed_point_vs_direct_solve(scale, seed):
    off, ids, w = _graph(scale, seed)
    n = len(off) - 1
    P, _ = _trans(off, ids, w)
    v = np.full(n, 1.0 / n)
    r, it, hist = oracle.wpagerank(off, ids, w, eps=1e-300, max_iter=400)
    assert it > 100 and hist[-1] < 1e-15  # run to a machine-precision fixed point
    assert _rel(r, _solve_r(P, v)) < 1e-12
    # personalised teleport (seeds only)
    vp = np.zeros(n)
    vp[[1, 7, n // 2]] = 1.0 / 3
    r2, _, _ = oracle.wpagerank(off, ids, w, eps=1e-300, max_iter=400, v=vp)
    assert _rel(r2, _solve_r(P, vp)) < 1e-12


def test_wpagerank_first_iterations_and_stop_rule():
    off, ids, w = _graph(9, 4)
    n = len(off) - 1
    P, _ = _trans(off, ids, w)
    v = np.full(n, 1.0 / n)
    r1, it, h = oracle.wpagerank(off, ids, w, eps=0.0, max_iter=1)
    want = (1 - ALPHA) * v + ALPHA * (P @ v)
    assert it == 1 and _rel(r1, want) < 1e-14
    assert np.isclose(h[0], np.abs(want - v).sum(), rtol=1e-12)
    # the stop rule: the first iteration whose L1 step is below eps
    r, it, h = oracle.wpagerank(off, ids, w, eps=1e-7, max_iter=100)
    assert h[-1] < 1e-7 and np.all(h[:-1] >= 1e-7)
    rk = v.copy()
    for _ in range(it):
        rk = (1 - ALPHA) * v + ALPHA * (P @ rk)
    assert _rel(r, rk) < 1e-13


def _loss_r(r, T):
    return float(np.sum(r[T] ** 2))


def test_wpagerank_adjoint_vs_direct_solve_and_euler():
    off, ids, w = _graph(11, 4)
    n = len(off) - 1
    P, S = _trans(off, ids, w)
    v = np.full(n, 1.0 / n)
    r, _, _ = oracle.wpagerank(off, ids, w, eps=1e-300, max_iter=400)
    T = np.random.default_rng(0).choice(n, 64, replace=False)
    u = np.zeros(n)
    u[T] = 2 * r[T]
    lam, gw, gv, it = oracle.wpagerank_adjoint(off, ids, w, r, u, eps=1e-300, max_iter=400)
    assert _rel(lam, _solve_lam(P, u)) < 1e-12
    assert np.array_equal(gv, (1 - ALPHA) * lam)
    # L is homogeneous of degree 2 in v (r is linear in v)
    assert np.isclose(np.dot(v, gv), 2 * _loss_r(r, T), rtol=1e-11)
    # P is invariant under scaling one source's out-weights: per source,
    # sum_e w_e dL/dw_e = 0
    src = _sources(off)
    per_src = np.bincount(src, weights=np.asarray(w, np.float64) * gw, minlength=n)
    assert np.abs(per_src).max() < 1e-12 * np.abs(gw).max() * 64


def test_wpagerank_adjoint_brute_force_tiny():
    """Dense brute force: dL/dw_e = alpha lambda^T (dP/dw_e) r with dP/dw_e
    written out column by column."""
    off, ids = oracle.csr_from_edges(7, [0, 0, 1, 2, 2, 3, 4, 5, 5, 6, 1], [1, 2, 2, 0, 3, 4, 5, 6, 0, 2, 5])
    n, m = 7, len(ids)
    w = np.array([0.5, 0.25, 1.0, 0.75, 0.125, 0.5, 0.375, 1.0, 0.625, 0.25, 0.875][:m], np.float32)
    P, S = _trans(off, ids, w)
    Pd = P.toarray()
    v = np.full(n, 1.0 / n)
    r = np.linalg.solve(np.eye(n) - ALPHA * Pd, (1 - ALPHA) * v)
    T = np.array([2, 5])
    u = np.zeros(n)
    u[T] = 2 * r[T]
    lam = np.linalg.solve(np.eye(n) - ALPHA * Pd.T, u)
    src = _sources(off)
    want = np.empty(m)
    for e in range(m):
        s, d = src[e], ids[e]
        dP = np.zeros((n, n))
        for e2 in range(off[s], off[s + 1]):
            dP[ids[e2], s] -= float(w[e2]) / S[s] ** 2
        dP[d, s] += 1.0 / S[s]
        want[e] = ALPHA * lam @ dP @ r
    ro, _, _ = oracle.wpagerank(off, ids, w, eps=1e-300, max_iter=600)
    lo, gw, gv, _ = oracle.wpagerank_adjoint(off, ids, w, ro, u, eps=1e-300, max_iter=600)
    assert np.allclose(ro, r, rtol=1e-13, atol=0)
    assert np.allclose(gw, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
    # and central differences on the exact solve
    for e in range(m):
        h = 1e-6
        wp, wm = w.astype(np.float64), w.astype(np.float64)
        wp[e] += h
        wm[e] -= h
        Pp = _trans(off, ids, wp)[0].toarray()
        Pm = _trans(off, ids, wm)[0].toarray()
        rp = np.linalg.solve(np.eye(n) - ALPHA * Pp, (1 - ALPHA) * v)
        rm = np.linalg.solve(np.eye(n) - ALPHA * Pm, (1 - ALPHA) * v)
        fd = (_loss_r(rp, T) - _loss_r(rm, T)) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-7 * abs(gw).max()


def test_wpagerank_adjoint_finite_differences_sampled():
    """FP64 central differences of L(w) through the exact fixed point on a
    scale-9 graph for sampled edges and teleport entries."""
    off, ids, w = _graph(9, 4)
    n, m = len(off) - 1, len(ids)
    v = np.full(n, 1.0 / n)
    P, _ = _trans(off, ids, w)
    r = _solve_r(P, v)
    T = np.random.default_rng(5).choice(n, 32, replace=False)
    u = np.zeros(n)
    u[T] = 2 * r[T]
    ro, _, _ = oracle.wpagerank(off, ids, w, eps=1e-300, max_iter=400)
    _, gw, gv, _ = oracle.wpagerank_adjoint(off, ids, w, ro, u, eps=1e-300, max_iter=400)
    rng = np.random.default_rng(6)
    wd = w.astype(np.float64)
    scale = np.abs(gw).max()
    for e in rng.choice(m, 16, replace=False):
        h = 1e-5
        wp, wm = wd.copy(), wd.copy()
        wp[e] += h
        wm[e] -= h
        fd = (_loss_r(_solve_r(_trans(off, ids, wp)[0], v), T) - _loss_r(_solve_r(_trans(off, ids, wm)[0], v), T)) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-6 * scale, (e, fd, gw[e])
    for i in rng.choice(n, 8, replace=False):
        h = 1e-6
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        fd = (_loss_r(_solve_r(P, vp), T) - _loss_r(_solve_r(P, vm), T)) / (2 * h)
        assert abs(fd - gv[i]) <= 1e-6 * np.abs(gv).max()


def test_wpagerank_adjoint_stop_rule():
    off, ids, w = _graph(10, 4)
    n = len(off) - 1
    r, _, _ = oracle.wpagerank(off, ids, w)
    u = np.zeros(n)
    u[:16] = 2 * r[:16]
    lam, _, _, it = oracle.wpagerank_adjoint(off, ids, w, r, u, eps=1e-7, max_iter=100)
    assert 1 < it < 100
    # zero seed: lambda stays zero and the first step satisfies the rule
    lam0, gw0, gv0, it0 = oracle.wpagerank_adjoint(off, ids, w, r, np.zeros(n))
    assert it0 == 1 and not lam0.any() and not gw0.any()
