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
    """cuda:0, after checking a GPU is present (GPU tests never skip)."""
    import torch
    assert torch.cuda.is_available(), "gpu-marked test run without a GPU"
    return torch.device("cuda:0")
