"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point declared in include/*.h, and rejects bad arguments with status
codes instead of aborting (no GPU is touched)."""
import ctypes as C
import glob
import os
import re

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
