"""The drop-in C++ front end (include/gg/ligra-numa.h).

CPU: the reference's four app sources compile unchanged against it (where
the reference tree exists).  GPU: a program written against the front end
(tests/cpp/frontend_parity.cpp) matches the oracle, and the reference apps
built against it run and report their timings.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import refcheck as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "apps", "_build")
REF_APP = "/root/reference/app"


@pytest.mark.skipif(not os.path.isdir(REF_APP), reason="reference sources not present")
@pytest.mark.parametrize("app", ["PageRank", "BFS", "Components", "SPMV"])
def test_reference_app_compiles_unchanged(app):
    with open(os.path.join(REF_APP, app + ".C")) as src:
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-", "-I.", "-I.."],
                           stdin=src, cwd=os.path.join(ROOT, "include", "gg"),
                           capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def _run(exe, app, args):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        env = dict(os.environ, GG_APP=app, GG_DUMP=out)
        r = subprocess.run([os.path.join(BUILD, exe)] + args + ["-rounds", "1"], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        raw = np.fromfile(out, np.float64)
        return raw[1:], int(raw[:1].view(np.int64)[0]), r.stdout


@pytest.mark.gpu
def test_frontend_pagerank():
    off, ids = oracle.rmat_csr(14, 16, 1)
    want, it, _ = oracle.Graph(off, ids, P=64).pagerank(10)
    got, iters, out = _run("frontend_parity", "pagerank", ["-c", "64", "-rmat", "14", "16", "1"])
    assert iters == it and R.rel_l1(got, want) < 1e-12
    assert "Average" in out


@pytest.mark.gpu
def test_frontend_bfs_and_cc():
    off, ids = oracle.rmat_csr(14, 16, 2, symmetric=True)
    got, levels, _ = _run("frontend_parity", "bfs", ["-s", "-c", "64", "-r", "100", "-rmat", "14", "16", "2"])
    R.check_parents(off, ids, 100, got.astype(np.int64))
    lab, rounds, _ = _run("frontend_parity", "cc", ["-s", "-c", "64", "-rmat", "14", "16", "2"])
    assert np.array_equal(lab.astype(np.int64), R.cc_min_labels(off, ids))


@pytest.mark.gpu
def test_frontend_spmv_weighted():
    off, ids = oracle.rmat_csr(12, 16, 3)
    w = oracle.weights(3, len(ids)).astype(np.float64)
    n = len(off) - 1
    want = R.to_scipy(off, ids, w).T @ np.full(n, 1.0 / n)
    got, _, _ = _run("frontend_parity_w", "spmv", ["-c", "64", "-rmat", "12", "16", "3"])
    assert R.rel_l1(got, want) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("app,args", [("PageRank", ["-c", "384", "-rmat", "18", "16", "1"]),
                                      ("BFS", ["-s", "-c", "384", "-r", "100", "-rmat", "18", "16", "2"]),
                                      ("Components", ["-s", "-c", "384", "-rmat", "18", "16", "2"]),
                                      ("SPMV", ["-c", "384", "-rmat", "18", "16", "3"])])
def test_reference_apps_run_on_gpu(app, args):
    exe = os.path.join(BUILD, app)
    if not os.path.exists(exe):
        pytest.skip("drop-in app binaries are built only where the reference sources exist")
    r = subprocess.run([exe] + args + ["-rounds", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("Running") == 2 and "Average" in r.stdout
