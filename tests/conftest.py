import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def dev():
    """The GPU the parity tests run on; fails loudly when there is none."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import hashgraph_amd
    hashgraph_amd.context(0)
    return torch.device("cuda", 0)
