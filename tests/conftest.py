import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests of the HIP kernels")
    config.addinivalue_line("markers", "slow: CPU test taking more than ~20 s")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip_device():
    """cuda:0 with libgratopy_hip.so loaded; fails (not skips) when a GPU test
    runs without a device or without the library."""
    import torch

    from gratopy_amd import _hip

    assert torch.cuda.is_available(), "GPU test needs a HIP device"
    _hip.load_library()
    return "cuda:0"
