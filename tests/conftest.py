import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "multi-block-lbm-gpu_amd")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long CPU run")


def gpu_available():
    """True when a HIP device is visible (checked through the library, not torch)."""
    try:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def lib32():
    import mbl

    return mbl.load_library(False)


@pytest.fixture(scope="session")
def lib64():
    import mbl

    return mbl.load_library(True)


@pytest.fixture(scope="session")
def olib():
    import oracle

    return oracle.load()


@pytest.fixture(autouse=True)
def _require_gpu(request):
    if request.node.get_closest_marker("gpu") and not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible")
