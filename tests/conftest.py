import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available():
    """A HIP device is visible to the runtime the engine links (libamdhip64
    from /opt/rocm), asked directly: torch bundles a runtime of its own, which
    may not see the device once the engine's runtime holds it."""
    import ctypes
    try:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    except OSError:
        return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


@pytest.fixture(scope="session")
def oracle():
    from oracle import hl_oracle
    hl_oracle.lib()
    return hl_oracle
