import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "opencfd-scu_amd"))
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libocfd_hip.so on the GPU)")


def gpu_available() -> bool:
    """Whether an AMD GPU is visible, asked of the kernel driver's topology
    without starting any HIP runtime in this process: the library's runtime
    must be the first to open the device (a runtime started here first, such
    as torch's own copy, leaves the library's hipSetDevice failing)."""
    import glob
    import os
    if not os.path.exists("/dev/kfd"):
        return False
    for p in glob.glob("/sys/class/kfd/kfd/topology/nodes/*/gpu_id"):
        try:
            if int(open(p).read().strip() or 0) != 0:
                return True
        except (OSError, ValueError):
            pass
    return False


@pytest.fixture(scope="session")
def hiplib():
    from pyocfd import capi
    return capi.load_library()
