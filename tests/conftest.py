import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible ROCm device; fails (never skips) without one."""
    import torch
    from evoxels_amd import _lib
    assert torch.cuda.is_available(), "GPU test run without a visible ROCm device"
    _lib.load()
    return torch.device("cuda", 0)
