"""firefly_amd -- MI355X (gfx950) wavefront path tracer.

The compute path is libff_hip.so (hand-written HIP kernels behind the C ABI
in include/ff/ff_abi.h).  Python here is only a thin ctypes layer used by
tests and benchmarks; there is no CPU fallback: loading fails loudly when
the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libff_hip.so")

_lib = None


def lib() -> ctypes.CDLL:
    """Load libff_hip.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m firefly_amd.build` "
                "(the gfx950 kernels are the only implementation)")
        from . import abi  # noqa: F401  (declares prototypes)
        _lib = abi.load(LIB_PATH)
    return _lib


__all__ = ["lib", "LIB_PATH"]
