"""ctypes binding of the C oracle (liboracle.so) and the reference harness.

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP product:
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  See gg_oracle.h for what each function restates and which
check pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PATH_SPARSE, PATH_CSC, PATH_COO = 0, 1, 2
OP_PR, OP_BFS, OP_CC, OP_SPMV = 0, 1, 2, 3

_lib = None


class Trace(C.Structure):
    _fields_ = [("path", C.c_int32), ("pad", C.c_int32), ("in_active", C.c_int64),
                ("in_out_edges", C.c_int64), ("out_active", C.c_int64),
                ("out_out_edges", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = np.ctypeslib.ndpointer
        u32 = P(np.uint32, flags="C_CONTIGUOUS")
        u64 = P(np.uint64, flags="C_CONTIGUOUS")
        f64 = P(np.float64, flags="C_CONTIGUOUS")
        i32 = P(np.int32, flags="C_CONTIGUOUS")
        L.ggo_splitmix64.restype = C.c_uint64
        L.ggo_splitmix64.argtypes = [C.c_uint64]
        L.ggo_rmat_raw.argtypes = [C.c_int, C.c_uint64, C.c_uint64, u32, u32]
        L.ggo_build_csr.restype = C.c_int64
        L.ggo_build_csr.argtypes = [C.c_uint32, C.c_uint64, u32, u32, C.c_int, u32, u32]
        L.ggo_transpose.argtypes = [C.c_uint32, C.c_uint64, u32, u32, u32, u32, u32]
        L.ggo_weight.restype = C.c_float
        L.ggo_weight.argtypes = [C.c_uint64, C.c_uint64]
        L.ggo_uniform_pm1.restype = C.c_double
        L.ggo_uniform_pm1.argtypes = [C.c_uint64, C.c_uint64]
        L.ggo_partition_sizes.argtypes = [C.c_uint64, u32, C.c_uint64, C.c_int, C.c_int,
                                          C.c_int, u64]
        L.ggo_hilbert_xy2d.restype = C.c_int64
        L.ggo_hilbert_xy2d.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        L.ggo_round_up_pow2.restype = C.c_uint64
        L.ggo_round_up_pow2.argtypes = [C.c_uint64]
        L.ggo_graph_new.restype = C.c_void_p
        L.ggo_graph_new.argtypes = [C.c_uint32, C.c_uint64, u32, u32, C.c_void_p, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int]
        L.ggo_graph_free.argtypes = [C.c_void_p]
        L.ggo_graph_starts.argtypes = [C.c_void_p, C.c_int, u64]
        L.ggo_graph_ncompressed.restype = C.c_uint64
        L.ggo_graph_ncompressed.argtypes = [C.c_void_p]
        L.ggo_graph_coo_count.restype = C.c_int64
        L.ggo_graph_coo_count.argtypes = [C.c_void_p, C.c_int]
        L.ggo_graph_coo.restype = C.c_int64
        L.ggo_graph_coo.argtypes = [C.c_void_p, C.c_int, u32, u32]
        L.ggo_pagerank.argtypes = [C.c_void_p, C.c_int, f64, C.POINTER(C.c_int), C.c_void_p]
        L.ggo_pagerank_ex.argtypes = [C.c_void_p, C.c_int, C.c_int, f64, C.POINTER(C.c_int), C.c_void_p]
        L.ggo_bfs.argtypes = [C.c_void_p, C.c_uint32, i32, C.c_void_p, C.c_int]
        L.ggo_cc.argtypes = [C.c_void_p, u32, C.c_void_p, C.c_int]
        L.ggo_spmv.argtypes = [C.c_void_p, C.c_void_p, f64, C.c_void_p]
        L.ggo_plus_reduce.restype = C.c_double
        L.ggo_plus_reduce.argtypes = [f64, C.c_uint64]
        L.ggo_edgemap_once.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, P(np.uint8, flags="C_CONTIGUOUS"),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int)]
        L.ggo_graph_error.restype = C.c_int
        f32 = P(np.float32, flags="C_CONTIGUOUS")
        L.ggo_out_wsum.argtypes = [C.c_uint32, u32, f32, f64]
        L.ggo_spmv_csr.argtypes = [C.c_uint32, u32, u32, C.c_void_p, f64, f64]
        L.ggo_spmv_bwd.argtypes = [C.c_uint32, u32, u32, C.c_void_p, f64, f64, C.c_void_p, C.c_void_p]
        L.ggo_wpagerank.argtypes = [C.c_uint32, u32, u32, f32, C.c_double, C.c_double, C.c_int,
                                    C.c_void_p, f64, C.c_void_p]
        L.ggo_wpagerank_adjoint.argtypes = [C.c_uint32, u32, u32, f32, C.c_double, C.c_double, C.c_int,
                                            f64, f64, f64, C.c_void_p, C.c_void_p]
        i64 = P(np.int64, flags="C_CONTIGUOUS")
        L.ggo_bc.argtypes = [C.c_uint32, u32, u32, C.c_uint32, f64, i32, C.c_void_p]
        L.ggo_sssp.argtypes = [C.c_uint32, u32, u32, f32, C.c_uint32, i64]
        L.ggo_prdelta.argtypes = [C.c_uint32, u32, u32, C.c_double, C.c_double, C.c_double, C.c_int, f64]
        L.ggo_prconv.argtypes = [C.c_uint32, u32, u32, C.c_double, C.c_double, C.c_int, f64]
        L.ggo_vebo.argtypes = [C.c_uint32, u32, C.c_int, u32, P(np.uint64, flags="C_CONTIGUOUS")]
        _lib = L
    return _lib


# ---------------------------------------------------------------- inputs --

def splitmix64(x: int) -> int:
    return lib().ggo_splitmix64(x)


def rmat_raw(scale: int, ef: int, seed: int):
    nraw = ef << scale
    src = np.empty(2 * nraw, np.uint32)
    dst = np.empty(2 * nraw, np.uint32)
    rc = lib().ggo_rmat_raw(scale, nraw, seed, src, dst)
    assert rc == 0
    return src, dst, nraw


def rmat_csr(scale: int, ef: int, seed: int, symmetric: bool = False):
    """R-MAT graph as CSR (offsets uint32 n+1, ids uint32 m)."""
    src, dst, nraw = rmat_raw(scale, ef, seed)
    n = 1 << scale
    off = np.empty(n + 1, np.uint32)
    ids = np.empty(2 * nraw if symmetric else nraw, np.uint32)
    m = lib().ggo_build_csr(n, nraw, src, dst, int(symmetric), off, ids)
    assert m >= 0, m
    return off, ids[:m].copy()


def csr_from_edges(n: int, src, dst, symmetric: bool = False):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    nraw = len(src)
    s2 = np.empty(2 * nraw + 1, np.uint32)
    d2 = np.empty(2 * nraw + 1, np.uint32)
    s2[:nraw] = src
    d2[:nraw] = dst
    off = np.empty(n + 1, np.uint32)
    ids = np.empty(2 * nraw + 1, np.uint32)
    m = lib().ggo_build_csr(n, nraw, s2, d2, int(symmetric), off, ids)
    assert m >= 0, m
    return off, ids[:m].copy()


def transpose(off, ids):
    n = len(off) - 1
    m = len(ids)
    toff = np.empty(n + 1, np.uint32)
    tids = np.empty(max(m, 1), np.uint32)
    perm = np.empty(max(m, 1), np.uint32)
    lib().ggo_transpose(n, m, off, ids, toff, tids, perm)
    return toff, tids[:m], perm[:m]


def weights(seed: int, m: int):
    L = lib()
    return np.array([L.ggo_weight(seed, e) for e in range(m)], np.float32)


def uniform_pm1(seed: int, n: int):
    L = lib()
    return np.array([L.ggo_uniform_pm1(seed, i) for i in range(n)], np.float64)


def partition_sizes(deg, m: int, P: int, mode: int = 0, relabel: int = 0):
    deg = np.ascontiguousarray(deg, np.uint32)
    sizes = np.zeros(P, np.uint64)
    rc = lib().ggo_partition_sizes(len(deg), deg, m, P, mode, relabel, sizes)
    return rc, sizes


def hilbert_xy2d(side: int, x: int, y: int) -> int:
    return lib().ggo_hilbert_xy2d(side, x, y)


def plus_reduce(a) -> float:
    a = np.ascontiguousarray(a, np.float64)
    return lib().ggo_plus_reduce(a, len(a))


# ---------------------------------------------------------------- engine --

class Graph:
    """Oracle graph: CSR + CSC twin + P partitions (+ Hilbert COO lists)."""

    def __init__(self, off, ids, w=None, symmetric=False, P=384, part_vertex=False,
                 hilbert=True, relabel=False):
        self.off = np.ascontiguousarray(off, np.uint32)
        self.ids = np.ascontiguousarray(ids, np.uint32)
        self.n = len(self.off) - 1
        self.m = len(self.ids)
        self.P = P
        self._w = None if w is None else np.ascontiguousarray(w, np.float64)
        L = lib()
        wp = None if self._w is None else self._w.ctypes.data
        self.h = L.ggo_graph_new(self.n, self.m, self.off, self.ids if self.m else
                                 np.zeros(1, np.uint32), wp, int(symmetric), P,
                                 int(part_vertex), int(hilbert), int(relabel))
        if not self.h:
            raise ValueError(f"oracle graph build failed: {L.ggo_graph_error()}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ggo_graph_free(self.h)
            self.h = None

    def starts(self, which: int):
        s = np.zeros(self.P + 1, np.uint64)
        lib().ggo_graph_starts(self.h, which, s)
        return s

    def ncompressed(self) -> int:
        return lib().ggo_graph_ncompressed(self.h)

    def coo(self, p: int):
        cnt = lib().ggo_graph_coo_count(self.h, p)
        if cnt < 0:
            return None
        s = np.empty(max(cnt, 1), np.uint32)
        d = np.empty(max(cnt, 1), np.uint32)
        lib().ggo_graph_coo(self.h, p, s, d)
        return s[:cnt], d[:cnt]

    def pagerank(self, max_iter=10, bit_frontiers=False):
        """PageRank.C (bit_frontiers: PageRankBit.C); (ranks, iters, trace)."""
        out = np.empty(self.n, np.float64)
        it = C.c_int(0)
        tr = (Trace * max_iter)()
        rc = lib().ggo_pagerank_ex(self.h, max_iter, int(bit_frontiers), out, C.byref(it),
                                   C.cast(tr, C.c_void_p))
        assert rc == 0, rc
        return out, it.value, [tr[i] for i in range(it.value)]

    def bfs(self, root: int, max_trace=256):
        par = np.empty(self.n, np.int32)
        tr = (Trace * max_trace)()
        lv = lib().ggo_bfs(self.h, root, par, C.cast(tr, C.c_void_p), max_trace)
        assert lv >= 0, lv
        return par, lv, [tr[i] for i in range(min(lv, max_trace))]

    def cc(self, max_trace=4096):
        ids = np.empty(self.n, np.uint32)
        tr = (Trace * max_trace)()
        r = lib().ggo_cc(self.h, ids, C.cast(tr, C.c_void_p), max_trace)
        assert r >= 0, r
        return ids, r, [tr[i] for i in range(min(r, max_trace))]

    def spmv(self, x=None):
        y = np.empty(self.n, np.float64)
        tr = Trace()
        xp = None
        if x is not None:
            x = np.ascontiguousarray(x, np.float64)
            xp = x.ctypes.data
        rc = lib().ggo_spmv(self.h, xp, y, C.byref(tr))
        assert rc == 0, rc
        return y, tr


# ------------------------------------------------------ reference harness --

# ------------------------------------------------- differentiated forms --

def _vp(a):
    return None if a is None else a.ctypes.data


def out_wsum(off, w):
    n = len(off) - 1
    S = np.empty(n, np.float64)
    lib().ggo_out_wsum(n, off, np.ascontiguousarray(w, np.float32), S)
    return S


def spmv_csr(off, ids, w, x):
    """y[d] = sum_{e=(s->d)} w_e x[s] in FP64 (w None: unweighted)."""
    n = len(off) - 1
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    y = np.empty(n, np.float64)
    lib().ggo_spmv_csr(n, off, ids, _vp(w), np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_bwd(off, ids, w, x, gy):
    """(gx, gw): gx[s] = sum_e w_e gy[d]; gw[e] = x[s] gy[d] in CSR order."""
    n, m = len(off) - 1, len(ids)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    gx = np.empty(n, np.float64)
    gw = np.empty(max(m, 1), np.float64)
    lib().ggo_spmv_bwd(n, off, ids, _vp(w), np.ascontiguousarray(x, np.float64),
                       np.ascontiguousarray(gy, np.float64), _vp(gx), _vp(gw))
    return gx, gw[:m]


def wpagerank(off, ids, w, alpha=0.85, eps=1e-7, max_iter=100, v=None):
    """Row-normalised weighted PageRank (all active); returns (r, iters, l1 history)."""
    n = len(off) - 1
    r = np.empty(n, np.float64)
    hist = np.zeros(max_iter, np.float64)
    v = None if v is None else np.ascontiguousarray(v, np.float64)
    it = lib().ggo_wpagerank(n, off, ids, np.ascontiguousarray(w, np.float32), alpha, eps, max_iter,
                             _vp(v), r, _vp(hist))
    assert it >= 0
    return r, it, hist[:it]


def wpagerank_adjoint(off, ids, w, r, u, alpha=0.85, eps=1e-7, max_iter=100):
    """Implicit adjoint at r for dL/dr = u; returns (lambda, gw (CSR order), gv, iters)."""
    n, m = len(off) - 1, len(ids)
    lam = np.empty(n, np.float64)
    gw = np.empty(max(m, 1), np.float64)
    gv = np.empty(n, np.float64)
    it = lib().ggo_wpagerank_adjoint(n, off, ids, np.ascontiguousarray(w, np.float32), alpha, eps, max_iter,
                                     np.ascontiguousarray(r, np.float64), np.ascontiguousarray(u, np.float64),
                                     lam, _vp(gw), _vp(gv))
    assert it >= 0
    return lam, gw[:m], gv, it


# ------------------------------------------------------- further apps --

def bc(off, ids, root):
    """app/BC.C from root: (score, level, sigma, rounds); score NaN where unreached."""
    n = len(off) - 1
    score = np.empty(n, np.float64)
    level = np.empty(n, np.int32)
    sigma = np.empty(n, np.float64)
    r = lib().ggo_bc(n, np.ascontiguousarray(off, np.uint32), np.ascontiguousarray(ids, np.uint32), root,
                     score, level, sigma.ctypes.data)
    assert r >= 0
    return score, level, sigma, r


def sssp(off, ids, w, root):
    """app/BellmanFord.C from root with integer lengths: (dist int64, rounds)."""
    n = len(off) - 1
    dist = np.empty(n, np.int64)
    r = lib().ggo_sssp(n, np.ascontiguousarray(off, np.uint32), np.ascontiguousarray(ids, np.uint32),
                       np.ascontiguousarray(w, np.float32), root, dist)
    assert r >= 0
    return dist, r


def prdelta(off, ids, damping=0.85, eps=1e-7, eps2=0.01, max_rounds=1 << 30):
    """app/PageRankDelta.C: (p, rounds)."""
    n = len(off) - 1
    p = np.empty(n, np.float64)
    r = lib().ggo_prdelta(n, np.ascontiguousarray(off, np.uint32), np.ascontiguousarray(ids, np.uint32),
                          damping, eps, eps2, max_rounds, p)
    assert r >= 0
    return p, r


def prconv(off, ids, damping=0.85, eps=1e-7, max_iter=100):
    """app/PageRankConverage.C: (p, iterations)."""
    n = len(off) - 1
    p = np.empty(n, np.float64)
    r = lib().ggo_prconv(n, np.ascontiguousarray(off, np.uint32), np.ascontiguousarray(ids, np.uint32),
                         damping, eps, max_iter, p)
    assert r >= 0
    return p, r


def vebo(indeg, P):
    """graphtools/VEBO.C: (new id of each vertex, partition starts) for P
    partitions."""
    n = len(indeg)
    out = np.empty(n, np.uint32)
    starts = np.empty(P + 1, np.uint64)
    r = lib().ggo_vebo(n, np.ascontiguousarray(indeg, np.uint32), P, out, starts)
    assert r == 0, r
    return out, starts


def ref_available(app: str = "pagerank") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"ref_{app}"))


def write_ggb(path, off, ids, w=None, symmetric=False):
    n = len(off) - 1
    m = len(ids)
    flags = (1 if symmetric else 0) | (2 if w is not None else 0)
    with open(path, "wb") as f:
        f.write(struct.pack("<4Q", 0x31424747, n, m, flags))
        f.write(np.ascontiguousarray(off, np.uint32).tobytes())
        f.write(np.ascontiguousarray(ids, np.uint32).tobytes())
        if w is not None:
            f.write(np.ascontiguousarray(w, np.int32).tobytes())


def run_ref(app: str, off, ids, w=None, symmetric=False, P=384, part_vertex=False,
            root=0, x=None, with_partitions=False, threads=None, relabel=False, trace=False):
    """Run oracle/_ref/ref_<app> (the reference engine) on a graph.

    relabel builds the partitioned graph with partition_relabel (the driver's
    -o); with trace the frontier's (d_m, num_out_edges) before each edgeMap
    is returned as an int64 array of pairs.  Returns (result float64[n],
    iters, partitions-or-None[, trace])."""
    exe = os.path.join(REF_DIR, f"ref_{app}")
    n = len(off) - 1
    with tempfile.TemporaryDirectory() as td:
        gp = os.path.join(td, "g.ggb")
        op = os.path.join(td, "out.bin")
        pp = os.path.join(td, "part.bin")
        write_ggb(gp, off, ids, w, symmetric)
        cmd = [exe, "-i", gp, "-o", op, "-c", str(P), "-v",
               "vertex" if part_vertex else "edge", "-r", str(root)]
        if with_partitions:
            cmd += ["-p", pp]
        if relabel:
            cmd += ["-R"]
        tp = os.path.join(td, "trace.bin")
        if trace:
            cmd += ["-t", tp]
        if x is not None:
            xp = os.path.join(td, "x.bin")
            np.ascontiguousarray(x, np.float64).tofile(xp)
            cmd += ["-x", xp]
        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"{exe} failed ({r.returncode}): {r.stderr[-2000:]}")
        raw = np.fromfile(op, np.float64)
        iters = int(raw[:1].view(np.int64)[0])
        res = raw[1:]
        parts = None
        if with_partitions:
            buf = np.fromfile(pp, np.uint8)
            o = 0
            coo_starts = buf[o:o + 8 * (P + 1)].view(np.int64).copy(); o += 8 * (P + 1)
            csc_starts = buf[o:o + 8 * (P + 1)].view(np.int64).copy(); o += 8 * (P + 1)
            lists = []
            if not part_vertex:
                for _ in range(P):
                    cnt = int(buf[o:o + 8].view(np.int64)[0]); o += 8
                    sd = buf[o:o + 8 * cnt].view(np.uint32).reshape(-1, 2).copy(); o += 8 * cnt
                    lists.append(sd)
            parts = (coo_starts, csc_starts, lists)
        assert len(res) == n
        if trace:
            return res, iters, parts, np.fromfile(tp, np.int64).reshape(-1, 2)
        return res, iters, parts
