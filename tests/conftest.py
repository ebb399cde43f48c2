import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    """A GPU is there when the ROCm driver's device node exists, or when the
    HIP library reports a device.  With the device node present the GPU tests
    are never skipped: a library that fails to load makes them fail."""
    if os.path.exists("/dev/kfd"):
        return True
    try:
        import phoebe_amd
        return phoebe_amd.device_count() > 0
    except Exception:  # pragma: no cover
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
