import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


FDTD_ERR_NODEVICE = -6


def gpu_available():
    """Ask the engine library itself (fdtd_device_name): a missing or broken
    library raises instead of skipping, and only the engine's own "no
    device" status (FDTD_ERR_NODEVICE) counts as no GPU."""
    import ctypes
    from wipems_amd._native import engine_lib
    lib = engine_lib()
    buf = ctypes.create_string_buffer(256)
    rc = lib.fdtd_device_name(buf, 256)
    if rc == FDTD_ERR_NODEVICE:
        return False
    if rc != 0:
        err = ctypes.create_string_buffer(1024)
        lib.fdtd_last_error(err, 1024)
        raise RuntimeError("fdtd_device_name failed (%d): %s" % (rc, err.value.decode(errors="replace")))
    return True


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU (fdtd_device_name: FDTD_ERR_NODEVICE)")
    return True
