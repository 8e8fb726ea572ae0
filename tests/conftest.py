import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: full-size (2^25-sample) GPU case")


@pytest.fixture(scope="session")
def device():
    """The GPU; a gpu-marked test fails (never skips) without one."""
    import torch
    assert torch.cuda.is_available(), "gpu test run without a visible GPU"
    return torch.device("cuda:0")
