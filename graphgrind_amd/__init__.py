"""graphgrind_amd -- MI355X-native GraphGrind traversal engine (host mirror).

Python view of the C ABI in include/graphgrind_amd.h.  Names follow the
reference's app-facing API (graphgrind/ligra-numa.h, graph-numa.h):

    Graph            partitioned_graph: CSR + in-edge twin + partitions on HBM
    Frontier         partitioned_vertices (create / bits / dense / sparse)
    edge_map         edgeMap(GA, frontier, functor, threshold, remDups)
    vertex_map       vertexMap(part, frontier, functor) / map_vertexL
    pagerank, bfs, components, spmv   the apps' Compute() loops
    spmv_bwd, wpagerank, wpagerank_adjoint   differentiated forms (adjoints)
    bc, sssp         app/BC.C and app/BellmanFord.C on the same engine
    msbfs, Graph.arena, autotune_pull        batching and tuning

Errors come back as GGError carrying the C status and message (the
reference abort()s in the same places).  The compute path is the HIP
library only: importing this package fails loudly when
libgraphgrind_amd.so is missing (build it with __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgraphgrind_amd.so")

GG_OK = 0
ERR_NAMES = {-1: "INVALID", -2: "HIP", -3: "OOM", -4: "SIZE", -5: "PARTITION",
             -6: "UNSUPPORTED", -7: "STATE"}
F64, F32, U32, I32, U8 = 0, 1, 2, 3, 4
NP_OF = {F64: np.float64, F32: np.float32, U32: np.uint32, I32: np.int32, U8: np.uint8}
DT_OF = {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.uint32): U32,
         np.dtype(np.int32): I32, np.dtype(np.uint8): U8}
OP_PAGERANK, OP_BFS, OP_CC, OP_SPMV, OP_SSSP, OP_PATHSUM, OP_SUM = 0, 1, 2, 3, 4, 5, 6
VOP_PAGERANK, VOP_RESET, VOP_CC_COPY, VOP_ABSDIFF = 0, 1, 2, 3
FRONTIER_EMPTY, FRONTIER_SINGLE, FRONTIER_ALL = 0, 1, 2
PATH_EMPTY, PATH_SPARSE, PATH_CSC, PATH_COO = -1, 0, 1, 2
PATH_NAMES = {-1: "empty", 0: "sparse", 1: "csc", 2: "coo"}
OPT_PULL_SCHEDULE, OPT_ROW_EDGES_PER_LANE, OPT_DENSE_COO, OPT_XCD_SEGMENTS = 0, 1, 2, 3


class GGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"GG_ERR_{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Op(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("a", C.c_void_p), ("b", C.c_void_p)]


class VOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("a", C.c_void_p), ("b", C.c_void_p),
                ("c0", C.c_double), ("c1", C.c_double)]


class GraphView(C.Structure):
    _fields_ = [("toff", C.c_void_p), ("tids", C.c_void_p), ("outdeg", C.c_void_p), ("perm", C.c_void_p),
                ("tile_row", C.c_void_p), ("tile_edge", C.c_void_p), ("tiles", C.c_uint32),
                ("n", C.c_uint32), ("m", C.c_uint64), ("row_work", C.c_void_p),
                ("row_blocks", C.c_uint32), ("row_long", C.c_uint32)]


class TuneInfo(C.Structure):
    _fields_ = [("schedule", C.c_int32), ("edges_per_lane", C.c_int32), ("ms", C.c_double),
                ("candidates", C.c_int32), ("cand_schedule", C.c_int32 * 8),
                ("cand_edges_per_lane", C.c_int32 * 8), ("cand_ms", C.c_double * 8)]


class RunInfo(C.Structure):
    _fields_ = [("iters", C.c_int32), ("paths", C.c_int32 * 64), ("edges_scanned", C.c_uint64),
                ("ms", C.c_double)]

    def path_list(self):
        return [PATH_NAMES[self.paths[i]] for i in range(min(self.iters, 64))]


_lib = None

# Every entry point of include/graphgrind_amd.h with (restype, argtypes).
_VP, _U64P, _U32P, _I32P, _F32P, _U8P = (C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                         C.POINTER(C.c_uint8))
_IP = C.POINTER(C.c_int)
SIGNATURES = {
    "gg_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gg_ctx_destroy": (C.c_int, [_VP]),
    "gg_ctx_sync": (C.c_int, [_VP]),
    "gg_last_error": (C.c_char_p, []),
    "gg_version": (C.c_char_p, []),
    "gg_malloc": (C.c_int, [_VP, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gg_free": (C.c_int, [_VP, _VP]),
    "gg_host_alloc": (C.c_int, [_VP, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gg_host_free": (C.c_int, [_VP, _VP]),
    "gg_memcpy_h2d": (C.c_int, [_VP, _VP, _VP, C.c_uint64]),
    "gg_memcpy_d2h": (C.c_int, [_VP, _VP, _VP, C.c_uint64]),
    "gg_memcpy_d2d": (C.c_int, [_VP, _VP, _VP, C.c_uint64]),
    "gg_fill": (C.c_int, [_VP, _VP, C.c_int, C.c_uint64, C.c_double]),
    "gg_iota": (C.c_int, [_VP, _VP, C.c_int, C.c_uint64]),
    "gg_graph_create_csr": (C.c_int, [_VP, C.c_uint32, C.c_uint64, _VP, _VP, _VP, C.c_int,
                                      C.POINTER(C.c_void_p)]),
    "gg_graph_create_rmat": (C.c_int, [_VP, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    "gg_graph_destroy": (C.c_int, [_VP]),
    "gg_graph_info": (C.c_int, [_VP, _U32P, _U64P, _IP, _IP]),
    "gg_graph_copy_csr": (C.c_int, [_VP, _VP, _VP, _VP]),
    "gg_graph_copy_csc": (C.c_int, [_VP, _VP, _VP, _VP]),
    "gg_graph_degrees": (C.c_int, [_VP, C.c_int, _VP]),
    "gg_graph_set_weights": (C.c_int, [_VP, _VP]),
    "gg_graph_partition": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "gg_graph_reorder": (C.c_int, [_VP, C.c_int]),
    "gg_graph_transpose": (C.c_int, [_VP]),
    "gg_vebo_order": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "gg_graph_relabel": (C.c_int, [_VP, _VP, C.POINTER(C.c_void_p)]),
    "gg_graph_device_view": (C.c_int, [_VP, C.c_int, C.POINTER(GraphView)]),
    "gg_graph_partition_starts": (C.c_int, [_VP, C.c_int, _VP, _IP]),
    "gg_graph_coo": (C.c_int, [_VP, _VP, _VP, _VP]),
    "gg_frontier_create": (C.c_int, [_VP, C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    "gg_frontier_from_host": (C.c_int, [_VP, _VP, _VP, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gg_frontier_destroy": (C.c_int, [_VP]),
    "gg_frontier_info": (C.c_int, [_VP, _U64P, _U64P, _IP, _IP, _U32P]),
    "gg_frontier_to_dense": (C.c_int, [_VP]),
    "gg_frontier_set_bit": (C.c_int, [_VP, C.c_int]),
    "gg_frontier_to_sparse": (C.c_int, [_VP]),
    "gg_frontier_copy_dense": (C.c_int, [_VP, _VP]),
    "gg_frontier_copy_sparse": (C.c_int, [_VP, _VP, C.c_uint64, _U64P]),
    "gg_edgemap": (C.c_int, [_VP, _VP, C.POINTER(Op), C.c_int64, C.c_int, C.POINTER(C.c_void_p),
                             _IP]),
    "gg_vertexmap": (C.c_int, [_VP, _VP, C.POINTER(VOp)]),
    "gg_vertexfilter": (C.c_int, [_VP, _VP, _VP, C.POINTER(C.c_void_p)]),
    "gg_reduce_sum": (C.c_int, [_VP, _VP, C.c_int, C.c_uint64, C.POINTER(C.c_double)]),
    "gg_pagerank": (C.c_int, [_VP, C.c_int, C.c_int, C.c_double, C.c_double, _VP,
                              C.POINTER(RunInfo)]),
    "gg_bfs": (C.c_int, [_VP, C.c_uint32, C.c_int64, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_cc": (C.c_int, [_VP, _VP, C.POINTER(RunInfo)]),
    "gg_spmv": (C.c_int, [_VP, C.c_int, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_spmv_noalloc": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "gg_last_kernel_time": (C.c_int, [_VP, C.POINTER(C.c_double), _IP]),
    "gg_graph_set_option": (C.c_int, [_VP, C.c_int, C.c_int64]),
    "gg_graph_get_option": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int64)]),
    "gg_graph_set_weights_dev": (C.c_int, [_VP, _VP]),
    "gg_graph_create_arena": (C.c_int, [_VP, C.c_int, _VP, C.POINTER(C.c_void_p), _VP, _VP]),
    "gg_msbfs": (C.c_int, [_VP, C.c_int, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_msbc_forward": (C.c_int, [_VP, C.c_int, _VP, C.c_int, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_autotune_pull": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(TuneInfo)]),
    "gg_bc": (C.c_int, [_VP, C.c_uint32, C.c_int, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_sssp": (C.c_int, [_VP, C.c_uint32, _VP, C.POINTER(RunInfo)]),
    "gg_spmv_bwd": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, _VP, C.POINTER(RunInfo)]),
    "gg_wpagerank": (C.c_int, [_VP, C.c_int, C.c_double, C.c_double, C.c_int, _VP, _VP,
                               C.POINTER(RunInfo)]),
    "gg_wpagerank_adjoint": (C.c_int, [_VP, C.c_int, C.c_double, C.c_double, C.c_int, _VP, _VP, _VP,
                                       _VP, _VP, C.POINTER(RunInfo)]),
}


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: the HIP extension is not built "
                              "(run __graft_entry__.build() or python -m graphgrind_amd.build_lib)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != GG_OK:
        raise GGError(rc, lib().gg_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data


def _dev(a, count=None, dtype=None, ctx=None, name="array"):
    """(device pointer, numpy dtype) of a DeviceArray, or of a contiguous
    device tensor exposing data_ptr() (a torch tensor on a HIP device; the
    caller orders its own stream before the call).

    With 'count' the array must hold at least that many elements, with
    'dtype' it must have that element type, and with 'ctx' it must live on
    that context's device.  A mismatch raises ValueError before any device
    work, since the library writes count * itemsize bytes through the
    pointer."""
    if a is None:
        return None, None
    if isinstance(a, DeviceArray):
        ptr, dt, size = a.ptr, a.dtype, a.count
        if ptr is None:
            raise ValueError(f"{name}: freed DeviceArray")
        if ctx is not None and a.ctx.device != ctx.device:
            raise ValueError(f"{name}: on device {a.ctx.device}, the graph is on device {ctx.device}")
    elif hasattr(a, "data_ptr"):
        if getattr(a.device, "type", None) != "cuda":
            raise ValueError(f"{name}: device tensor required")
        if not a.is_contiguous():
            raise ValueError(f"{name}: contiguous tensor required")
        if ctx is not None and a.device.index is not None and a.device.index != ctx.device:
            raise ValueError(f"{name}: tensor on device {a.device.index}, the graph is on device {ctx.device}")
        ptr, dt, size = a.data_ptr(), np.dtype(str(a.dtype).replace("torch.", "")), a.numel()
    else:
        raise TypeError(f"{name}: not a device array: {type(a)!r}")
    if dtype is not None and dt != np.dtype(dtype):
        raise ValueError(f"{name}: dtype {dt}, expected {np.dtype(dtype)}")
    if count is not None and size < count:
        raise ValueError(f"{name}: {size} elements, at least {count} required")
    return ptr, dt


# ----------------------------------------------------------------- memory --

class Context:
    """One GPU and one HIP stream (replaces the Cilk/NUMA runtime setup)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().gg_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if self.h:
            lib().gg_ctx_destroy(self.h)
            self.h = None

    def sync(self):
        check(lib().gg_ctx_sync(self.h))

    def kernel_time(self):
        ms = C.c_double()
        k = C.c_int()
        check(lib().gg_last_kernel_time(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value


class DeviceArray:
    """mmap_ptr<T>: an HBM allocation of 'count' elements."""

    def __init__(self, ctx: Context, count: int, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().gg_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        check(lib().gg_memcpy_h2d(ctx.h, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.count, self.dtype)
        check(lib().gg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def fill(self, value):
        check(lib().gg_fill(self.ctx.h, self.ptr, DT_OF[self.dtype], self.count, float(value)))
        return self

    def iota(self):
        check(lib().gg_iota(self.ctx.h, self.ptr, DT_OF[self.dtype], self.count))
        return self

    def free(self):
        if self.ptr and self.ctx.h:  # a closed context already released the device
            lib().gg_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ graph --

class Graph:
    """partitioned_graph on the device: CSR, in-edge twin, partitions."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.h = handle
        n = C.c_uint32()
        m = C.c_uint64()
        s = C.c_int()
        w = C.c_int()
        check(lib().gg_graph_info(self.h, C.byref(n), C.byref(m), C.byref(s), C.byref(w)))
        self.n, self.m, self.symmetric, self.weighted = n.value, m.value, bool(s.value), bool(w.value)
        self.P = 0

    @classmethod
    def from_csr(cls, ctx, off, ids, w=None, symmetric=False):
        off = np.ascontiguousarray(off, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        wv = None if w is None else np.ascontiguousarray(w, np.float32)
        h = C.c_void_p()
        check(lib().gg_graph_create_csr(ctx.h, len(off) - 1, len(ids), off.ctypes.data,
                                        ids.ctypes.data if len(ids) else None, _ptr(wv),
                                        int(symmetric), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def rmat(cls, ctx, scale, edge_factor=16, seed=1, symmetric=False, weighted=False):
        h = C.c_void_p()
        check(lib().gg_graph_create_rmat(ctx.h, scale, edge_factor, seed, int(symmetric),
                                         int(weighted), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def arena(cls, ctx, parts):
        """Packs graphs into one block-diagonal arena graph.  Returns (arena,
        vbase, ebase): graph i owns vertices [vbase[i], vbase[i+1]) and edges
        [ebase[i], ebase[i+1]) of the arena."""
        hs = (C.c_void_p * len(parts))(*[p.h for p in parts])
        vb = np.zeros(len(parts) + 1, np.uint32)
        eb = np.zeros(len(parts) + 1, np.uint64)
        h = C.c_void_p()
        check(lib().gg_graph_create_arena(ctx.h, len(parts), hs, C.byref(h), vb.ctypes.data, eb.ctypes.data))
        return cls(ctx, h), vb.astype(np.int64), eb.astype(np.int64)

    def close(self):
        if self.h and self.ctx.h:
            lib().gg_graph_destroy(self.h)
        self.h = None

    def partition(self, P=384, by_vertex=False, relabel=False, hilbert=True):
        """-c P, -v edge|vertex, -o, EDGES_HILBERT (ligra-numa.h:1104-1124)."""
        check(lib().gg_graph_partition(self.h, P, int(by_vertex), int(relabel), int(hilbert)))
        self.P = P
        self.part_vertex = by_vertex
        return self

    def transpose(self):
        """GA.transpose() (graph-numa.h:547-555, 721-725): later calls run on
        the reversed edges; a second call restores the graph."""
        check(lib().gg_graph_transpose(self.h))
        return self

    def vebo_order(self, P):
        """graphtools/VEBO.C order for P partitions: (new id of each vertex,
        first new id of each partition), numpy arrays."""
        new = np.empty(self.n, np.uint32)
        starts = np.empty(P + 1, np.uint64)
        check(lib().gg_vebo_order(self.h, P, new.ctypes.data, starts.ctypes.data))
        return new, starts

    def relabel(self, new_of_old):
        """A new graph with vertex v renumbered new_of_old[v]."""
        perm = np.ascontiguousarray(new_of_old, np.uint32)
        if len(perm) != self.n:
            raise ValueError("relabel: one new id per vertex")
        h = C.c_void_p()
        check(lib().gg_graph_relabel(self.h, perm.ctypes.data, C.byref(h)))
        return Graph(self.ctx, h)

    def reorder(self, mode=1):
        """Internal hot-sources-first renumbering used by the pull kernels."""
        check(lib().gg_graph_reorder(self.h, mode))
        return self

    def device_view(self, which=0):
        v = GraphView()
        check(lib().gg_graph_device_view(self.h, which, C.byref(v)))
        return v

    def starts(self, which=0):
        s = np.zeros(self.P + 1, np.uint64)
        u = C.c_int()
        check(lib().gg_graph_partition_starts(self.h, which, s.ctypes.data, C.byref(u)))
        return s, bool(u.value)

    def coo(self):
        src = np.empty(max(self.m, 1), np.uint32)
        dst = np.empty(max(self.m, 1), np.uint32)
        po = np.zeros(self.P + 1, np.uint64)
        check(lib().gg_graph_coo(self.h, src.ctypes.data, dst.ctypes.data, po.ctypes.data))
        return src[:self.m], dst[:self.m], po

    def csr(self):
        off = np.empty(self.n + 1, np.uint32)
        ids = np.empty(max(self.m, 1), np.uint32)
        w = np.empty(max(self.m, 1), np.float32) if self.weighted else None
        check(lib().gg_graph_copy_csr(self.h, off.ctypes.data, ids.ctypes.data, _ptr(w)))
        return off, ids[:self.m], (None if w is None else w[:self.m])

    def csc(self):
        off = np.empty(self.n + 1, np.uint32)
        ids = np.empty(max(self.m, 1), np.uint32)
        w = np.empty(max(self.m, 1), np.float32) if self.weighted else None
        check(lib().gg_graph_copy_csc(self.h, off.ctypes.data, ids.ctypes.data, _ptr(w)))
        return off, ids[:self.m], (None if w is None else w[:self.m])

    def degrees(self, in_edges=False):
        d = np.empty(self.n, np.uint32)
        check(lib().gg_graph_degrees(self.h, int(in_edges), d.ctypes.data))
        return d

    def set_option(self, option, value):
        """Schedule choices of the device pull (OPT_PULL_SCHEDULE, OPT_ROW_EDGES_PER_LANE)."""
        check(lib().gg_graph_set_option(self.h, option, int(value)))
        return self

    def get_option(self, option):
        v = C.c_int64()
        check(lib().gg_graph_get_option(self.h, option, C.byref(v)))
        return v.value

    def set_weights_dev(self, w):
        """Edge weights from a device array (m FP32 values, CSR order)."""
        p, dt = _dev(w, self.m, np.float32, self.ctx, "weights")
        check(lib().gg_graph_set_weights_dev(self.h, p))
        self.weighted = True

    def set_weights(self, w_csr):
        w = np.ascontiguousarray(w_csr, np.float32)
        if w.size != self.m:
            raise ValueError(f"set_weights: {w.size} weights for {self.m} edges")
        check(lib().gg_graph_set_weights(self.h, w.ctypes.data))
        self.weighted = True


class Frontier:
    """partitioned_vertices (ligra-numa.h:91-272)."""

    def __init__(self, graph: Graph, handle):
        self.g = graph
        self.h = handle

    @classmethod
    def create(cls, g, kind, v=0):
        h = C.c_void_p()
        check(lib().gg_frontier_create(g.h, kind, v, C.byref(h)))
        return cls(g, h)

    @classmethod
    def single(cls, g, v):
        return cls.create(g, FRONTIER_SINGLE, v)

    @classmethod
    def bits(cls, g):
        return cls.create(g, FRONTIER_ALL)

    @classmethod
    def empty(cls, g):
        return cls.create(g, FRONTIER_EMPTY)

    @classmethod
    def from_dense(cls, g, flags):
        f = np.ascontiguousarray(flags, np.uint8)
        h = C.c_void_p()
        check(lib().gg_frontier_from_host(g.h, f.ctypes.data, None, len(f), C.byref(h)))
        return cls(g, h)

    @classmethod
    def from_sparse(cls, g, ids):
        a = np.ascontiguousarray(ids, np.uint32)
        h = C.c_void_p()
        check(lib().gg_frontier_from_host(g.h, None, a.ctypes.data if len(a) else None, len(a),
                                          C.byref(h)))
        return cls(g, h)

    def info(self):
        dm = C.c_uint64()
        oe = C.c_uint64()
        dense = C.c_int()
        bit = C.c_int()
        n = C.c_uint32()
        check(lib().gg_frontier_info(self.h, C.byref(dm), C.byref(oe), C.byref(dense),
                                     C.byref(bit), C.byref(n)))
        return {"d_m": dm.value, "out_edges": oe.value, "dense": bool(dense.value),
                "bit": bool(bit.value), "n": n.value}

    def numNonzeros(self):
        return self.info()["d_m"]

    def isEmpty(self):
        return self.numNonzeros() == 0

    def set_bit(self, bit=True):
        """The public 'bit' field (all vertices active, stored d_m and
        num_out_edges kept), as app/PageRankBit.C:145 sets it."""
        check(lib().gg_frontier_set_bit(self.h, int(bool(bit))))

    def to_dense(self):
        check(lib().gg_frontier_to_dense(self.h))

    def to_sparse(self):
        check(lib().gg_frontier_to_sparse(self.h))

    def dense(self):
        out = np.zeros(self.g.n, np.uint8)
        check(lib().gg_frontier_copy_dense(self.h, out.ctypes.data))
        return out

    def sparse(self):
        cap = self.g.n + self.info()["d_m"] + 1
        out = np.empty(cap, np.uint32)
        cnt = C.c_uint64()
        check(lib().gg_frontier_copy_sparse(self.h, out.ctypes.data, cap, C.byref(cnt)))
        return out[:cnt.value]

    def close(self):
        if self.h and self.g.ctx.h:
            lib().gg_frontier_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# element types of each edge operator's arrays a and b ("float": FP64 or
# FP32, b the same as a), as include/graphgrind_amd.h documents them
_EDGE_OP_TYPES = {OP_PAGERANK: ("float", "float"), OP_BFS: (np.int32, None), OP_CC: (np.uint32, np.uint32),
                  OP_SPMV: ("float", "float"), OP_SSSP: (np.int32, np.int32), OP_PATHSUM: ("float", np.uint8),
                  OP_SUM: ("float", "float")}


def edge_map(g: Graph, frontier: Frontier, kind, a: DeviceArray, b: DeviceArray = None,
             threshold=-1, remdups=False):
    """edgeMap(GA, frontier, F, threshold, DENSE, remDups) -> (output, path)."""
    ap, adt = _dev(a, g.n, None, g.ctx, "a")
    want_a, want_b = _EDGE_OP_TYPES.get(kind, (None, None))
    if want_a == "float":
        if adt not in (np.float64, np.float32):
            raise ValueError(f"a: FP64 or FP32 for operator {kind}, not {adt}")
    elif want_a is not None and adt != want_a:
        raise ValueError(f"a: dtype {adt}, operator {kind} takes {np.dtype(want_a)}")
    bp, _ = _dev(b, g.n, adt if want_b == "float" else want_b, g.ctx, "b")
    dtype = DT_OF[adt]
    op = Op(kind, dtype, ap, bp)
    h = C.c_void_p()
    path = C.c_int()
    check(lib().gg_edgemap(g.h, frontier.h, C.byref(op), threshold, int(remdups), C.byref(h),
                           C.byref(path)))
    return Frontier(g, h), path.value


def vertex_map(g: Graph, frontier, kind, a: DeviceArray, b: DeviceArray = None, c0=0.0, c1=0.0):
    """vertexMap(part, frontier, F); frontier None = map_vertexL over all."""
    ap, adt = _dev(a, g.n, None, g.ctx, "a")
    if adt not in DT_OF:
        raise ValueError(f"a: unsupported dtype {adt}")
    bp, _ = _dev(b, g.n, adt, g.ctx, "b")
    op = VOp(kind, DT_OF[adt], ap, bp, c0, c1)
    check(lib().gg_vertexmap(g.h, None if frontier is None else frontier.h, C.byref(op)))


def vertex_filter(g: Graph, frontier, flags):
    """vertexFilter(GA, V, F) (ligra-numa.h:1053-1082) with the filter's
    answers given as device flags (uint8, n): the members of 'frontier'
    whose flag is non-zero, as a new dense frontier."""
    fp, _ = _dev(flags, g.n, np.uint8, g.ctx, "flags")
    h = C.c_void_p()
    check(lib().gg_vertexfilter(g.h, frontier.h, fp, C.byref(h)))
    return Frontier(g, h)


def reduce_sum(ctx: Context, a: DeviceArray):
    out = C.c_double()
    ap, adt = _dev(a, None, None, ctx, "a")
    if adt not in DT_OF:
        raise ValueError(f"a: unsupported dtype {adt}")
    check(lib().gg_reduce_sum(ctx.h, ap, DT_OF[adt], a.count if isinstance(a, DeviceArray) else a.numel(),
                              C.byref(out)))
    return out.value


# ------------------------------------------------------------------- apps --

def pagerank(g: Graph, dtype=np.float64, max_iter=10, damping=0.85, eps=1e-7, out=None):
    """app/PageRank.C Compute(): returns (ranks DeviceArray, RunInfo)."""
    dt = np.dtype(dtype)
    if dt not in (np.float64, np.float32):
        raise ValueError(f"pagerank: FP64 or FP32, not {dt}")
    r = out if out is not None else DeviceArray(g.ctx, g.n, dt)
    rp, _ = _dev(r, g.n, dt, g.ctx, "out")
    info = RunInfo()
    check(lib().gg_pagerank(g.h, DT_OF[dt], max_iter, damping, eps, rp, C.byref(info)))
    return r, info


def bfs(g: Graph, root, threshold=-1, levels=True, parents=None, level_out=None):
    """app/BFS.C Compute(): returns (parents, levels|None, RunInfo)."""
    # the caller's arrays are checked before anything is allocated
    _dev(parents, g.n, np.int32, g.ctx, "parents")
    _dev(level_out, g.n, np.int32, g.ctx, "levels")
    p = parents if parents is not None else DeviceArray(g.ctx, g.n, np.int32)
    lv = level_out if level_out is not None else (DeviceArray(g.ctx, g.n, np.int32) if levels else None)
    pp, _ = _dev(p, g.n, np.int32, g.ctx, "parents")
    lp, _ = _dev(lv, g.n, np.int32, g.ctx, "levels")
    info = RunInfo()
    check(lib().gg_bfs(g.h, root, threshold, pp, lp, C.byref(info)))
    return p, lv, info


def components(g: Graph, out=None):
    """app/Components.C Compute(): returns (labels, RunInfo)."""
    lab = out if out is not None else DeviceArray(g.ctx, g.n, np.uint32)
    lp, _ = _dev(lab, g.n, np.uint32, g.ctx, "labels")
    info = RunInfo()
    check(lib().gg_cc(g.h, lp, C.byref(info)))
    return lab, info


def spmv(g: Graph, x, y=None):
    """y[d] = sum_{s->d} w * x[s] over the in-edge twin (app/SPMV.C).  x and y
    are DeviceArrays or device tensors."""
    xp, dt = _dev(x, g.n, None, g.ctx, "x")
    if dt not in (np.float64, np.float32):
        raise ValueError(f"x: FP64 or FP32, not {dt}")
    yy = y if y is not None else DeviceArray(g.ctx, g.n, dt)
    yp, _ = _dev(yy, g.n, dt, g.ctx, "y")
    info = RunInfo()
    check(lib().gg_spmv(g.h, DT_OF[dt], xp, yp, C.byref(info)))
    return yy, info


# ---------------------------------------------------- differentiated forms --

def spmv_bwd(g: Graph, x, gy, gx=None, gw=None, want_gx=True, want_gw=True):
    """Backward of spmv for dL/dy = gy: gx[s] = sum_e w_e gy[d] (pull over
    the out-edge CSR), gw[e] = x[s] gy[d] in CSR order.  Arrays are
    DeviceArrays or device tensors of one float dtype; returns (gx, gw, info)."""
    gp, dt = _dev(gy, g.n, None, g.ctx, "gy")
    if dt not in (np.float64, np.float32):
        raise ValueError(f"gy: FP64 or FP32, not {dt}")
    xp, _ = _dev(x, g.n, dt, g.ctx, "x")
    if want_gx and gx is None:
        gx = DeviceArray(g.ctx, g.n, dt)
    if want_gw and gw is None:
        gw = DeviceArray(g.ctx, max(g.m, 1), dt)
    gxp = _dev(gx, g.n, dt, g.ctx, "gx")[0] if want_gx else None
    gwp = _dev(gw, g.m, dt, g.ctx, "gw")[0] if want_gw else None
    info = RunInfo()
    check(lib().gg_spmv_bwd(g.h, DT_OF[dt], xp, gp, gxp, gwp, C.byref(info)))
    return (gx if want_gx else None), (gw if want_gw else None), info


def wpagerank(g: Graph, dtype=np.float64, alpha=0.85, eps=1e-7, max_iter=100, v=None, out=None):
    """Row-normalised weighted PageRank with all vertices active (teleport v,
    uniform when None) until the L1 step is below eps; returns (r, info)."""
    dt = np.dtype(dtype)
    if dt not in (np.float64, np.float32):
        raise ValueError(f"wpagerank: FP64 or FP32, not {dt}")
    r = out if out is not None else DeviceArray(g.ctx, g.n, dt)
    vp, _ = _dev(v, g.n, dt, g.ctx, "v")
    rp, _ = _dev(r, g.n, dt, g.ctx, "out")
    info = RunInfo()
    check(lib().gg_wpagerank(g.h, DT_OF[dt], alpha, eps, max_iter, vp, rp, C.byref(info)))
    return r, info


def wpagerank_adjoint(g: Graph, r, u, alpha=0.85, eps=1e-7, max_iter=100, gw=None, gv=None, lam=None,
                      want_gw=True, want_gv=True, want_lam=False):
    """Implicit adjoint of wpagerank at r for dL/dr = u.  Returns
    (gw per edge in CSR order, gv = dL/dv, lambda, info)."""
    rp, dt = _dev(r, g.n, None, g.ctx, "r")
    if dt not in (np.float64, np.float32):
        raise ValueError(f"r: FP64 or FP32, not {dt}")
    up, _ = _dev(u, g.n, dt, g.ctx, "u")
    if want_gw and gw is None:
        gw = DeviceArray(g.ctx, max(g.m, 1), dt)
    if want_gv and gv is None:
        gv = DeviceArray(g.ctx, g.n, dt)
    if want_lam and lam is None:
        lam = DeviceArray(g.ctx, g.n, dt)
    gwp = _dev(gw, g.m, dt, g.ctx, "gw")[0] if want_gw else None
    gvp = _dev(gv, g.n, dt, g.ctx, "gv")[0] if want_gv else None
    lamp = _dev(lam, g.n, dt, g.ctx, "lam")[0] if want_lam else None
    info = RunInfo()
    check(lib().gg_wpagerank_adjoint(g.h, DT_OF[dt], alpha, eps, max_iter, rp, up, gwp, gvp, lamp,
                                     C.byref(info)))
    return (gw if want_gw else None), (gv if want_gv else None), (lam if want_lam else None), info


# ------------------------------------------------------- further apps --

def bc(g: Graph, root, dtype=np.float32, out=None, levels=True):
    """app/BC.C from root: (scores, levels|None, RunInfo); NaN where unreached."""
    sc = out if out is not None else DeviceArray(g.ctx, g.n, dtype)
    lv = DeviceArray(g.ctx, g.n, np.int32) if levels else None
    info = RunInfo()
    p, dt = _dev(sc, g.n, None, g.ctx, "out")
    if dt not in (np.float64, np.float32):
        raise ValueError(f"bc: FP64 or FP32, not {dt}")
    check(lib().gg_bc(g.h, root, DT_OF[dt], p, None if lv is None else lv.ptr, C.byref(info)))
    return sc, lv, info


def sssp(g: Graph, root, out=None):
    """app/BellmanFord.C from root with integer lengths: (dist int32, RunInfo)."""
    d = out if out is not None else DeviceArray(g.ctx, g.n, np.int32)
    info = RunInfo()
    check(lib().gg_sssp(g.h, root, _dev(d, g.n, np.int32, g.ctx, "out")[0], C.byref(info)))
    return d, info


# -------------------------------------------------- batching and tuning --

def msbfs(g: Graph, sources, out=None):
    """BFS from up to 32 sources at once (the forward phase of app/BC.C).
    Returns (levels: int32 DeviceArray [len(sources) * n], -1 unreached,
    RunInfo)."""
    src = np.ascontiguousarray(sources, np.uint32)
    lv = out if out is not None else DeviceArray(g.ctx, len(src) * g.n, np.int32)
    info = RunInfo()
    lvp, _ = _dev(lv, len(src) * g.n, np.int32, g.ctx, "out")
    check(lib().gg_msbfs(g.h, len(src), src.ctypes.data, lvp, C.byref(info)))
    return lv, info


def msbc_forward(g: Graph, sources, dtype=np.float64, paths=None, levels=False):
    """The forward phase of app/BC.C for up to 32 sources at once: returns
    (paths DeviceArray [len(sources) * n] of dtype, NumPaths of every vertex
    from each source, 0 unreached; levels int32 [len(sources) * n] or None;
    RunInfo)."""
    src = np.ascontiguousarray(sources, np.uint32)
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("msbc_forward: dtype must be float64 or float32")
    total = len(src) * g.n
    pa = paths if paths is not None else DeviceArray(g.ctx, total, dt)
    pp, _ = _dev(pa, total, dt, g.ctx, "paths")
    lv = DeviceArray(g.ctx, total, np.int32) if levels else None
    lvp = _dev(lv, total, np.int32, g.ctx, "levels")[0] if lv is not None else None
    info = RunInfo()
    check(lib().gg_msbc_forward(g.h, len(src), src.ctypes.data, DT_OF[dt], pp, lvp, C.byref(info)))
    return pa, lv, info


def autotune_pull(g: Graph, dtype=np.float64, trials=5):
    """Times the candidate pull schedules on g, keeps the fastest on the
    graph and returns every candidate's time."""
    t = TuneInfo()
    check(lib().gg_autotune_pull(g.h, DT_OF[np.dtype(dtype)], trials, C.byref(t)))
    k = t.candidates
    return {"schedule": t.schedule, "edges_per_lane": t.edges_per_lane, "ms": t.ms, "candidates": k,
            "cand_schedule": list(t.cand_schedule[:k]), "cand_edges_per_lane": list(t.cand_edges_per_lane[:k]),
            "cand_ms": list(t.cand_ms[:k])}
