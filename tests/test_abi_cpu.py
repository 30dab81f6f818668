"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, validates arguments before touching the device, and the
host-only SAH builder works."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from firefly_amd import abi, lib, scenes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ff", "ff_abi.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ff_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = lib()
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(abi.EXPORTED)
    assert L.ff_abi_version() == 2


def test_null_arguments_are_rejected():
    L = lib()
    assert L.ff_context_destroy(None) == 1
    assert L.ff_scene_destroy(None) == 1
    assert L.ff_envmap_destroy(None) == 1
    assert L.ff_render(None, None, None, None, None, None, None) == 1
    assert L.ff_trace_closest(None, None, 0, None, 0, None, None) == 1
    assert L.ff_trace_closest_noalloc(None, None, 0, None, 0, None, None) == 1
    assert L.ff_trace_any(None, None, 0, None, None, None) == 1
    assert L.ff_film_resolve(None, 0, 0, 1.0, None, None, None) == 1
    assert b"NULL" in L.ff_last_error() or b"bad" in L.ff_last_error()
    assert L.ff_bvh_build_gpu(None, None, 0, 8, None, None, None, None, None) == 1
    assert b"ff_bvh_build_gpu" in L.ff_last_error()
    with pytest.raises(abi.FFError) as e:
        abi.check(L.ff_scene_info(None, None, None, None))
    assert e.value.status == 1


def test_bvh_build_host_validates_and_builds():
    L = lib()
    n = C.c_int32()
    assert L.ff_bvh_build_host(None, 0, 8, None, None, C.byref(n), None) == 1
    pos = scenes.spheres_scene(segments=8, rings=4).positions
    nodes, ordered, depth = abi.bvh_build_host(pos)
    assert sorted(ordered.tolist()) == list(range(len(pos)))
    assert depth >= 2 and len(nodes) <= 2 * len(pos) - 1


def test_context_without_gpu_reports_error_or_succeeds():
    """Without a visible device ff_context_create must fail cleanly."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    h = C.c_void_p()
    st = lib().ff_context_create(0, C.byref(h))
    assert st in (1, 2) and not h.value
