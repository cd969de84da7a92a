import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu_device():
    """One lt_device for the whole GPU session (tests run in one process)."""
    import lantern_amd
    if not _has_gpu():
        pytest.fail("no GPU visible: GPU tests must run on a gfx950 box")
    dev = lantern_amd.Device(0)
    yield dev
    dev.close()
