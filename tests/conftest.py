import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "sdpx.jl_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

os.environ.setdefault("OMP_NUM_THREADS", "8")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded once; skips (never falls back) when no GPU is visible."""
    import ctypes

    from sdpx_mi355x import _lib

    lib = _lib.load()
    n = ctypes.c_int(0)
    lib.sdpx_device_count(ctypes.byref(n))
    if n.value < 1:
        pytest.fail("no GPU visible to the HIP library")
    return lib
