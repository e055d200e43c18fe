import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcnhip.so")
    config.addinivalue_line("markers", "slow: longer-running test")


@pytest.fixture
def rng(request):
    """Per-test RNG seeded from the test node id (as the reference's tests/conftest.py:24-29)."""
    import zlib
    return np.random.default_rng(zlib.crc32(request.node.nodeid.encode()))


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _gpu_guard(request):
    if request.node.get_closest_marker("gpu") is not None and not gpu_available():
        pytest.skip("no HIP device visible")
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
