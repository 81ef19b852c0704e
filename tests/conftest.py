import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libptyhip.so")


@pytest.fixture(scope="session")
def gpu_device():
    import torch
    from ptychography_amd import _abi
    _abi.load()
    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a GPU")
    return torch.device("cuda", 0)
