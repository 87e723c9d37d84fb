import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def gpu_present():
    try:
        import sadcs_amd
    except ImportError:
        return False
    status, _ = sadcs_amd.gpu_query_device()
    return status == sadcs_amd.GpuStatus.Ok


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a device; GPU tests fail (not skip) when the device or library is missing."""
    import sadcs_amd
    status, device = sadcs_amd.gpu_query_device()
    assert status == sadcs_amd.GpuStatus.Ok, f"no usable GPU: {sadcs_amd.gpu_last_error()}"
    print(f"device: {device.name.decode()} ({device.free_bytes >> 20} MiB free)", flush=True)
    return sadcs_amd
