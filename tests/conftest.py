import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libgpe_hip.so")


def gpu_available():
    """True when a HIP device is visible (checked without initialising HIP in
    this process when possible)."""
    if os.path.exists("/dev/kfd"):
        return True
    return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.skip("no HIP GPU")
