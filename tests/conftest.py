"""Shared test setup: markers, library loading, oracle access."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Device context for GPU tests; fails loudly (no fallback) if the HIP
    library is missing on a machine with a GPU."""
    if not _gpu_available():
        pytest.skip("no GPU visible")
    import torch
    from firefly_amd import abi, lib
    lib()  # raises if libff_hip.so is not built
    torch.cuda.init()
    ctx = abi.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
