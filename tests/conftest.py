import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def ovx():
    """The HIP backend, initialised on device 0 (GPU tests only)."""
    import torch
    import openvsip_amd
    assert torch.cuda.is_available(), "GPU test without a GPU"
    openvsip_amd.init(0)
    return openvsip_amd
