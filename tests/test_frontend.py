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
@pytest.mark.parametrize("app", ["PageRank", "BFS", "Components", "SPMV", "BC", "BellmanFord", "PageRankDelta",
                                 "PageRankBit", "PageRankConverage"])
def test_reference_app_compiles_unchanged(app):
    with open(os.path.join(REF_APP, app + ".C")) as src:
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-", "-I.", "-I.."],
                           stdin=src, cwd=os.path.join(ROOT, "include", "gg"),
                           capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_host_atomics_under_threads():
    """CAS / writeMin / writeAdd of the front end (utils.h:423-477) on float,
    int, unsigned, double and long from 8 OpenMP threads: exact sums and
    minima, and the neighbouring elements untouched (including a target at
    the end of its array)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "apps"), "_build/atomics"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([os.path.join(BUILD, "atomics")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 5, r.stdout


def _build_functor_check(tmp, app=None):
    """tests/cpp/functor_check.cpp, alone or behind a reference app source
    fed to the compiler unchanged (as apps/Makefile builds the apps)."""
    exe = os.path.join(tmp, "fc_" + (app or "neg"))
    src = open(os.path.join(ROOT, "tests", "cpp", "functor_check.cpp")).read()
    flags = []
    if app and app.startswith("frontend_parity"):
        src = open(os.path.join(ROOT, "tests", "cpp", "frontend_parity.cpp")).read() + "\n" + src
        flags = ["-DGG_FRONT_NO_DRIVER", "-DCHECK_FRONTEND_PARITY"] + (["-DWEIGHTED=1"] if app.endswith("_w") else [])
    elif app:
        src = open(os.path.join(REF_APP, app + ".C")).read() + "\n" + src
        flags = ["-DGG_FRONT_NO_DRIVER", "-DCHECK_" + app.upper()]
        if app in ("SPMV", "BellmanFord"):
            flags.append("-DWEIGHTED")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", "-w"] + flags +
                       ["-x", "c++", "-", "-I.", "-I..", "-o", exe,
                        "-L" + os.path.join(ROOT, "graphgrind_amd"), "-lgraphgrind_amd",
                        "-Wl,-rpath," + os.path.join(ROOT, "graphgrind_amd"),
                        "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"],
                       input=src, cwd=os.path.join(ROOT, "include", "gg"), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_functor_self_check_refuses_other_bodies(tmp_path):
    """Functors with PR_F / BFS_F / CC_F / PR_Vertex_Reset members but other
    update bodies are refused by the front end's self-check (edgeMap stops
    with a message, vertexMap runs them on the host)."""
    r = subprocess.run([_build_functor_check(str(tmp_path))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("refused") == 4, r.stdout


@pytest.mark.parametrize("prog", ["frontend_parity", "frontend_parity_w"])
def test_functor_self_check_accepts_parity_program_functors(tmp_path, prog):
    """The functors of this repository's GPU parity program pass the
    self-check (they run on the GPU in the -m gpu tests)."""
    r = subprocess.run([_build_functor_check(str(tmp_path), prog)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refused" not in r.stdout and "agrees" in r.stdout, r.stdout


@pytest.mark.skipif(not os.path.isdir(REF_APP), reason="reference sources not present")
@pytest.mark.parametrize("app", ["PageRank", "BFS", "Components", "SPMV", "BC", "BellmanFord", "PageRankDelta",
                                 "PageRankConverage"])
def test_functor_self_check_accepts_reference_functors(tmp_path, app):
    """The reference apps' own functors (their sources compiled unchanged)
    pass the self-check of the device operator they are mapped to."""
    r = subprocess.run([_build_functor_check(str(tmp_path), app)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refused" not in r.stdout and "agrees" in r.stdout, r.stdout


def _run(exe, app, args):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        env = dict(os.environ, GG_APP=app, GG_DUMP=out)
        # the input file is the last argument (commandLine::getArgument(0))
        r = subprocess.run([os.path.join(BUILD, exe), "-rounds", "1"] + args, env=env,
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
                                      ("SPMV", ["-c", "384", "-rmat", "18", "16", "3"]),
                                      ("BC", ["-c", "384", "-r", "100", "-rmat", "16", "16", "2"]),
                                      ("BellmanFord", ["-s", "-c", "384", "-r", "100", "-rmat", "16", "16", "2"]),
                                      ("PageRankDelta", ["-c", "384", "-rmat", "16", "16", "1"]),
                                      ("PageRankBit", ["-c", "384", "-rmat", "16", "16", "1"]),
                                      ("PageRankConverage", ["-c", "384", "-rmat", "16", "16", "1"])])
def test_reference_apps_run_on_gpu(app, args):
    exe = os.path.join(BUILD, app)
    if not os.path.exists(exe):
        pytest.skip("drop-in app binaries are built only where the reference sources exist")
    r = subprocess.run([exe] + args + ["-rounds", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("Running") == 2 and "Average" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("sym", [False, True])
def test_frontend_bc(sym):
    """BC.C's loop (both phases, host vertexMap functors, GA.transpose())
    through the front end against the oracle (pinned by Brandes):
    FP64 <= 1e-12 rel L1, the reference's FP32 fType <= 1e-4, NaN exactly
    where unreached."""
    off, ids = oracle.rmat_csr(12, 16, 2, symmetric=sym)
    want, lv, _, rounds = oracle.bc(off, ids, 100)
    reached = lv >= 0
    args = (["-s"] if sym else []) + ["-c", "64", "-r", "100", "-rmat", "12", "16", "2"]
    for app, tol in (("bc", 1e-12), ("bc32", 1e-4)):
        got, r, _ = _run("frontend_parity", app, args)
        assert r == rounds
        assert np.all(np.isnan(got[~reached])) and not np.any(np.isnan(got[reached]))
        assert R.rel_l1(got[reached], want[reached]) < tol, app


@pytest.mark.gpu
def test_frontend_bellman_ford(tmp_path):
    """BellmanFord.C's loop through the front end: unit lengths (unweighted
    build, as the reference without WEIGHTED) and integer lengths read from a
    PBBS WeightedAdjacencyGraph file; distances bit-exact."""
    off, ids = oracle.rmat_csr(12, 8, 3)
    want1, _ = oracle.sssp(off, ids, np.ones(len(ids), np.float32), 100)
    got, _, _ = _run("frontend_parity", "sssp", ["-c", "64", "-r", "100", "-rmat", "12", "8", "3"])
    assert np.array_equal(got.astype(np.int64), want1)
    w = (1 + (np.arange(len(ids)) * 2654435761 % 23)).astype(np.int64)
    path = str(tmp_path / "w.adj")
    write_pbbs(path, off, ids, w)
    want, _ = oracle.sssp(off, ids, w.astype(np.float32), 100)
    got, _, _ = _run("frontend_parity_w", "sssp", ["-c", "64", "-r", "100", path])
    assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.gpu
def test_frontend_pagerank_delta():
    """PageRankDelta.C's loop (signed pushes from a filtered frontier, host
    vertexFilter functors) through the front end against the oracle (pinned
    by power iteration): same round count, ranks <= 1e-10 rel L1."""
    off, ids = oracle.rmat_csr(13, 16, 1)
    want, rounds = oracle.prdelta(off, ids)
    got, r, _ = _run("frontend_parity", "prdelta", ["-c", "64", "-rmat", "13", "16", "1"])
    assert r == rounds
    assert R.rel_l1(got, want) < 1e-10


@pytest.mark.gpu
def test_frontend_pagerank_converage():
    """PageRankConverage.C's loop (edgeMap of add_factor, host Csum / norm
    functors over map_partition sums) through the front end against the
    oracle (pinned by a closed-form first step and a direct solve)."""
    off, ids = oracle.rmat_csr(13, 16, 1)
    want, it = oracle.prconv(off, ids)
    got, r, _ = _run("frontend_parity", "prconv", ["-c", "64", "-rmat", "13", "16", "1"])
    assert r == it
    assert R.rel_l1(got, want) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("app", ["prdelta", "prconv"])
def test_frontend_delta_and_converage_match_reference_golden(tmp_path, app):
    """PageRankDelta.C / PageRankConverage.C loops through the front end,
    reading the fixture's graph from a PBBS file, against the reference
    engine's own run on it (tests/golden/apps_*.npz)."""
    for path in R.app_golden_cases(app):
        z = R.load_golden(path)
        f = str(tmp_path / "g.adj")
        write_pbbs(f, z["off"], z["ids"])
        got, it, _ = _run("frontend_parity", app, ["-c", str(int(z["P"])), f])
        assert it == int(z[f"{app}_iters"]), path
        assert R.rel_l1(got, z[app]) < 1e-12, path


@pytest.mark.gpu
def test_frontend_frontier_factories():
    """dense/sparse (empty, n rows), boolean (every third vertex) and
    indice (vertices 0..9) frontiers from the front end."""
    off, ids = oracle.rmat_csr(10, 8, 1)
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    got, ok, _ = _run("frontend_parity", "factories", ["-c", "16", "-rmat", "10", "8", "1"])
    assert ok == 1
    assert list(got[:4]) == [0, 0, 0, 0]
    third = np.arange(0, n, 3)
    assert got[4] == len(third) and got[5] == deg[third].sum()
    assert got[6] == 10 and got[7] == deg[:10].sum()


# ---- input formats (IO.h:136-361 PBBS text, 365-656 Galois .gr) -------------

def write_pbbs(path, off, ids, w=None):
    """PBBS AdjacencyGraph / WeightedAdjacencyGraph text: header, n, m, the n
    row offsets, the m ids (then the m integer weights)."""
    n = len(off) - 1
    with open(path, "w") as f:
        f.write("WeightedAdjacencyGraph\n" if w is not None else "AdjacencyGraph\n")
        f.write(f"{n}\n{len(ids)}\n")
        f.write("\n".join(str(int(x)) for x in off[:-1]) + "\n")
        f.write("\n".join(str(int(x)) for x in ids) + "\n")
        if w is not None:
            f.write("\n".join(str(int(x)) for x in w) + "\n")


def write_galois(path, off, ids, w=None):
    """Galois .gr v1: u64 {1, edge-data size, n, m}, u64 row ends, u32 ids
    (padded to an even count), u32 edge data."""
    n, m = len(off) - 1, len(ids)
    with open(path, "wb") as f:
        np.array([1, 4 if w is not None else 0, n, m], np.uint64).tofile(f)
        np.asarray(off[1:], np.uint64).tofile(f)
        np.asarray(ids, np.uint32).tofile(f)
        if m % 2:
            np.zeros(1, np.uint32).tofile(f)
        if w is not None:
            np.asarray(w, np.uint32).tofile(f)


def _read(path, binary=False, weighted=False):
    exe = os.path.join(BUILD, "read_graph")
    if not os.path.exists(exe):
        pytest.skip("apps not built")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "csr.bin")
        args = [exe] + (["-b"] if binary else []) + (["-w"] if weighted else []) + [path, out]
        r = subprocess.run(args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (r.stdout, r.stderr)
        raw = open(out, "rb").read()
    n, m = np.frombuffer(raw[:16], np.uint64)
    n, m = int(n), int(m)
    off = np.frombuffer(raw[16:16 + 4 * (n + 1)], np.uint32)
    ids = np.frombuffer(raw[16 + 4 * (n + 1):16 + 4 * (n + 1 + m)], np.uint32)
    w = np.frombuffer(raw[16 + 4 * (n + 1 + m):], np.float32) if weighted else None
    return off, ids, w


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("binary", [False, True])
def test_readers_roundtrip(tmp_path, binary, weighted):
    off, ids = oracle.rmat_csr(9, 8, 11)
    w = (1 + np.arange(len(ids)) % 19).astype(np.int64) if weighted else None
    path = str(tmp_path / ("g.gr" if binary else "g.adj"))
    (write_galois if binary else write_pbbs)(path, off, ids, w)
    got_off, got_ids, got_w = _read(path, binary, weighted)
    assert np.array_equal(got_off, off) and np.array_equal(got_ids, ids)
    if weighted:
        assert np.array_equal(got_w, w.astype(np.float32))


def test_reader_odd_edge_count_padding(tmp_path):
    off, ids = oracle.csr_from_edges(4, [0, 1, 2], [1, 2, 3])  # m = 3: ids padded
    w = np.array([5, 6, 7])
    path = str(tmp_path / "odd.gr")
    write_galois(path, off, ids, w)
    o, i, ww = _read(path, True, True)
    assert np.array_equal(o, off) and np.array_equal(i, ids) and np.array_equal(ww, w.astype(np.float32))


@pytest.mark.skipif(not os.path.exists("/root/reference/sixnode"), reason="reference tree not present")
def test_reader_reference_sixnode():
    """The reference's own sample graph file parses to the CSR the golden
    fixtures were built from."""
    off, ids, _ = _read("/root/reference/sixnode")
    g = R.load_golden(os.path.join(ROOT, "tests", "golden", "sixnode_p2.npz"))
    assert np.array_equal(off, g["off"]) and np.array_equal(ids, g["ids"])


def test_reader_rejects_wrong_header(tmp_path):
    path = str(tmp_path / "bad.adj")
    with open(path, "w") as f:
        f.write("EdgeArray\n1\n0\n0\n")
    exe = os.path.join(BUILD, "read_graph")
    if not os.path.exists(exe):
        pytest.skip("apps not built")
    r = subprocess.run([exe, path, str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Bad input file" in r.stdout


@pytest.mark.gpu
def test_frontend_pagerank_from_pbbs_file(tmp_path):
    """The driver reading a PBBS file (no -rmat) gives the oracle's ranks."""
    off, ids = oracle.rmat_csr(12, 16, 4)
    path = str(tmp_path / "g.adj")
    write_pbbs(path, off, ids)
    want, it, _ = oracle.Graph(off, ids, P=64).pagerank(10)
    got, iters, _ = _run("frontend_parity", "pagerank", ["-c", "64", path])
    assert iters == it and R.rel_l1(got, want) < 1e-12


@pytest.mark.gpu
def test_frontend_bfs_from_galois_file(tmp_path):
    off, ids = oracle.rmat_csr(12, 16, 2, symmetric=True)
    path = str(tmp_path / "g.gr")
    write_galois(path, off, ids)
    got, _, _ = _run("frontend_parity", "bfs", ["-s", "-b", "-c", "64", "-r", "100", path])
    R.check_parents(off, ids, 100, got.astype(np.int64))
