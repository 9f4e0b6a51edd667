"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point declared in include/*.h, and rejects bad arguments with status
codes instead of aborting (no GPU is touched)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import graphgrind_amd as gg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(gg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gg_edgemap" in names and "gg_pagerank" in names and len(names) > 30


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    L = gg.lib()
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_python_mirror_binds_every_entry_point():
    assert set(declared_functions()) <= set(gg.SIGNATURES), \
        set(declared_functions()) - set(gg.SIGNATURES)


def test_argument_errors_without_gpu():
    L = gg.lib()
    assert L.gg_ctx_create(0, None) == -1
    assert b"NULL" in L.gg_last_error()
    assert L.gg_graph_info(None, None, None, None, None) == -1
    assert L.gg_frontier_destroy(None) == -1
    assert L.gg_edgemap(None, None, None, -1, 0, None, None) == -1
    assert L.gg_pagerank(None, 0, 10, 0.85, 1e-7, None, None) == -1
    assert L.gg_version().startswith(b"graphgrind_amd")


def test_product_never_loads_the_oracle():
    # the HIP path must not route through the checker
    for f in glob.glob(os.path.join(ROOT, "graphgrind_amd", "**", "*"), recursive=True):
        if f.endswith((".py", ".hip", ".h", ".cpp")):
            text = open(f).read()
            for needle in ("import oracle", "liboracle", "gg_oracle", "ggo_"):
                assert needle not in text, f"{f} references the oracle ({needle})"


class _Ctx:
    def __init__(self, device):
        self.device = device
        self.h = None


def _fake_array(ctx, count, dtype):
    """A DeviceArray record with a made-up address: never passed to the
    library, only to the wrappers' argument checks."""
    a = object.__new__(gg.DeviceArray)
    a.ctx, a.count, a.dtype, a.ptr = ctx, count, np.dtype(dtype), 0x1000
    a.nbytes = count * a.dtype.itemsize
    return a


class _FakeTensor:
    class _Dev:
        def __init__(self, t, i):
            self.type, self.index = t, i

    def __init__(self, n, dtype="float64", dev=("cuda", 0), contiguous=True):
        self.n, self.dtype, self.device, self.c = n, "torch." + dtype, self._Dev(*dev), contiguous

    def data_ptr(self):
        return 0x2000

    def numel(self):
        return self.n

    def is_contiguous(self):
        return self.c


def test_wrapper_argument_checks():
    """The Python wrappers refuse arrays that are too short, of the wrong
    element type or on another device before calling into the library (the
    library writes count * itemsize bytes through the pointer)."""
    import numpy as np  # noqa: F811

    ctx = _Ctx(0)
    a = _fake_array(ctx, 100, np.float64)
    assert gg._dev(a, 100, np.float64, ctx)[1] == np.float64
    with pytest.raises(ValueError, match="at least 101"):
        gg._dev(a, 101, np.float64, ctx)
    with pytest.raises(ValueError, match="dtype"):
        gg._dev(a, 100, np.float32, ctx)
    with pytest.raises(ValueError, match="device"):
        gg._dev(_fake_array(_Ctx(1), 100, np.float64), 100, np.float64, ctx)
    t = _FakeTensor(100, "float32")
    assert gg._dev(t, 100, np.float32, ctx)[1] == np.float32
    with pytest.raises(ValueError, match="at least"):
        gg._dev(t, 200, np.float32, ctx)
    with pytest.raises(ValueError, match="dtype"):
        gg._dev(t, 100, np.float64, ctx)
    with pytest.raises(ValueError, match="device 1"):
        gg._dev(_FakeTensor(100, "float32", ("cuda", 1)), 100, np.float32, ctx)
    with pytest.raises(ValueError, match="device tensor"):
        gg._dev(_FakeTensor(100, "float32", ("cpu", None)), 100, np.float32, ctx)
    with pytest.raises(ValueError, match="contiguous"):
        gg._dev(_FakeTensor(100, "float32", contiguous=False), 100, np.float32, ctx)
    # the app wrappers run the same checks before any library call
    g = object.__new__(gg.Graph)
    g.ctx, g.n, g.m, g.h = ctx, 100, 1000, None
    with pytest.raises(ValueError, match="dtype float32, expected float64"):
        gg.pagerank(g, np.float64, out=_fake_array(ctx, 100, np.float32))
    with pytest.raises(ValueError, match="at least 100"):
        gg.bfs(g, 0, parents=_fake_array(ctx, 99, np.int32))
    with pytest.raises(ValueError, match="labels"):
        gg.components(g, out=_fake_array(ctx, 100, np.int32))
    with pytest.raises(ValueError, match="y: dtype"):
        gg.spmv(g, _fake_array(ctx, 100, np.float32), _fake_array(ctx, 100, np.float64))
    with pytest.raises(ValueError, match="gw: 999 elements"):
        gg.spmv_bwd(g, _fake_array(ctx, 100, np.float32), _fake_array(ctx, 100, np.float32),
                    gx=_fake_array(ctx, 100, np.float32), gw=_fake_array(ctx, 999, np.float32))
    with pytest.raises(ValueError, match="at least 3200"):
        gg.msbfs(g, list(range(32)), out=_fake_array(ctx, 3199, np.int32))
    with pytest.raises(ValueError, match="b: dtype"):
        gg.edge_map(g, None, gg.OP_PATHSUM, _fake_array(ctx, 100, np.float64), _fake_array(ctx, 100, np.int32))
    with pytest.raises(ValueError, match="operator"):
        gg.edge_map(g, None, gg.OP_CC, _fake_array(ctx, 100, np.int32), _fake_array(ctx, 100, np.int32))
    with pytest.raises(ValueError, match="gv"):
        gg.wpagerank_adjoint(g, _fake_array(ctx, 100, np.float64), _fake_array(ctx, 100, np.float64),
                             gw=_fake_array(ctx, 1000, np.float64), gv=_fake_array(ctx, 10, np.float64))
