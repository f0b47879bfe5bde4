import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libencore_rq.so on cuda:0)")


@pytest.fixture(scope="session")
def device():
    """The engine's device handle on GPU 0 (GPU tests only)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no GPU is visible")
    from encore_amd import rq

    torch.cuda.init()
    dev = rq.Device(0)
    yield dev
    dev.close()
