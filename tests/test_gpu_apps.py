"""GPU parity of the apps beyond the four (BC.C, BellmanFord.C) against the
oracle (oracle/gg_oracle_apps.c, itself pinned by textbook Brandes and
Dijkstra).

Bars: BC levels bit-exact, scores <= 1e-12 rel L1 in FP64 and <= 1e-4 in
FP32 (the reference's own fType) over reached vertices, NaN exactly where
unreached; Bellman-Ford distances bit-exact (integers).
"""
import numpy as np
import pytest

import graphgrind_amd as gg
import oracle

pytestmark = pytest.mark.gpu

INF = 2147483647 // 2


@pytest.fixture(scope="module")
def ctx():
    c = gg.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale,ef,seed,sym,root", [(8, 8, 2, True, 3), (14, 16, 2, True, 100),
                                                    (12, 8, 5, False, 1), (16, 16, 1, False, 7),
                                                    (14, 16, 2, False, 100)])  # deepest level without out-edges
def test_bc_vs_oracle(ctx, scale, ef, seed, sym, root, plan):
    """plan 1: both phases on the hot-sources-first renumbering, scores and
    levels written back in the caller's numbering."""
    print(f"bc scale={scale} sym={sym} plan={plan}", flush=True)
    off, ids = oracle.rmat_csr(scale, ef, seed, symmetric=sym)
    want, wl, _, rounds = oracle.bc(off, ids, root)
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=sym).reorder(plan)
    reached = wl >= 0
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-4)):
        sc, lv, info = gg.bc(g, root, dt)
        got = sc.to_host().astype(np.float64)
        assert np.array_equal(lv.to_host(), wl)
        assert info.iters == rounds
        assert np.all(np.isnan(got[~reached])) and not np.any(np.isnan(got[reached]))
        assert _rel(got[reached], want[reached]) < tol, (dt, _rel(got[reached], want[reached]))
    g.close()


def test_bc_root_only_and_errors(ctx):
    off, ids = oracle.csr_from_edges(4, [1, 2], [2, 3])
    g = gg.Graph.from_csr(ctx, off, ids)
    sc, lv, info = gg.bc(g, 0, np.float64)
    s = sc.to_host()
    assert s[0] == 0.0 and np.all(np.isnan(s[1:])) and list(lv.to_host()) == [0, -1, -1, -1]
    with pytest.raises(gg.GGError):
        gg.bc(g, 4)
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
@pytest.mark.parametrize("scale,seed,sym", [(10, 3, False), (14, 4, True), (16, 3, False)])
def test_sssp_vs_oracle(ctx, scale, seed, sym, plan):
    print(f"sssp scale={scale} plan={plan}", flush=True)
    off, ids = oracle.rmat_csr(scale, 8, seed, symmetric=sym)
    w = (1 + (np.arange(len(ids)) * 2654435761 % 19)).astype(np.float32)
    want, _ = oracle.sssp(off, ids, w, 0)
    g = gg.Graph.from_csr(ctx, off, ids, w=w, symmetric=sym).reorder(plan)
    d, info = gg.sssp(g, 0)
    assert np.array_equal(d.to_host().astype(np.int64), want)
    assert info.iters >= 1
    g.close()
    # unit lengths when the graph has no weights
    g = gg.Graph.from_csr(ctx, off, ids, symmetric=sym).reorder(plan)
    want1, _ = oracle.sssp(off, ids, np.ones(len(ids), np.float32), 0)
    d1, _ = gg.sssp(g, 0)
    assert np.array_equal(d1.to_host().astype(np.int64), want1)
    g.close()


@pytest.mark.parametrize("plan", [0, 1])
def test_sssp_negative_cycle(ctx, plan):
    off, ids = oracle.csr_from_edges(3, [0, 1, 2], [1, 2, 1])
    w = np.array([1, 1, -3], np.float32)
    g = gg.Graph.from_csr(ctx, off, ids, w=w).reorder(plan)
    d, info = gg.sssp(g, 0)
    assert np.all(d.to_host() == -INF) and info.iters == 3
    g.close()


# ---- against the reference engine's own runs (tests/golden/apps_*.npz) ----

import refcheck as R  # noqa: E402


@pytest.mark.parametrize("path", R.app_golden_cases("bc"), ids=lambda p: p.split("/")[-1])
def test_bc_matches_reference_golden(ctx, path):
    """The reference's BC.C computes in FP32; both of our precisions must
    agree with it to FP32 rounding, with NaN exactly where it has NaN."""
    z = R.load_golden(path)
    want = z["bc"]
    ok = ~np.isnan(want)
    g = gg.Graph.from_csr(ctx, z["off"], z["ids"], symmetric=bool(z["symmetric"]))
    for dt, plan in ((np.float32, 0), (np.float64, 0), (np.float32, 1), (np.float64, 1)):
        g.reorder(plan)
        sc, _, info = gg.bc(g, int(z["root"]), dt)
        got = sc.to_host().astype(np.float64)
        assert info.iters == int(z["bc_iters"])
        assert np.array_equal(np.isnan(got), ~ok)
        assert _rel(got[ok], want[ok]) < 1e-5, dt
    g.close()


@pytest.mark.parametrize("path", R.app_golden_cases("bellmanford"), ids=lambda p: p.split("/")[-1])
def test_sssp_matches_reference_golden(ctx, path):
    z = R.load_golden(path)
    g = gg.Graph.from_csr(ctx, z["off"], z["ids"], w=z["bellmanford_w"].astype(np.float32),
                          symmetric=bool(z["symmetric"]))
    for plan in (0, 1):
        g.reorder(plan)
        d, _ = gg.sssp(g, int(z["root"]))
        assert np.array_equal(d.to_host().astype(np.int64), z["bellmanford"].astype(np.int64)), plan
    g.close()
