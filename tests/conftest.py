import os
import sys

import numpy as np
import pytest

# PyTorch-ROCm bundles its own HIP runtime (libamdhip64.so.7, the same soname as ROCm's).
# libwithin_amd.so binds to whichever copy the process has loaded first; when the library comes
# first and torch later, torch maps its second copy and cannot initialise the GPU.  The GPU
# tests that hand torch device buffers to the library therefore need torch loaded first.
try:
    import torch  # noqa: F401
except ImportError:  # the CPU suite does not need it
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import OracleLib
    return OracleLib()


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when the HIP library or the GPU is missing: GPU tests must
    exercise the native path."""
    from within_amd import _lib
    n, name = _lib.device_info()
    assert n > 0, "no GPU visible to the HIP runtime"
    print(f"[gpu] {name}", flush=True)
    return name


def cfg1_data(n=100_000, seed=1):
    """BASELINE config 1 / README quickstart data (README.md:29-45)."""
    np.random.seed(seed)
    fe = np.asfortranarray(np.column_stack([np.random.randint(0, 500, n),
                                            np.random.randint(0, 200, n)]).astype(np.uint32))
    y = np.random.randn(n)
    return fe, y
