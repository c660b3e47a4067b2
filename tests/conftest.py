import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "opencfd-scu_amd"))
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libocfd_hip.so on the GPU)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to ask whether a device is visible)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hiplib():
    from pyocfd import capi
    return capi.load_library()
