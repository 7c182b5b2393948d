import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpds.so on the device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available():
    try:
        import ctypes
        from parallel_ds_amd._native import lib
        n = ctypes.c_int(0)
        hip = ctypes.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
            return False
        lib()
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly on a machine that should have a GPU but cannot load the HIP library."""
    from parallel_ds_amd._native import lib
    lib()
    if not gpu_available():
        pytest.fail("no usable GPU for a test marked gpu")
    return True
