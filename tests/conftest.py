import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libflygpu.so")


@pytest.fixture(scope="session")
def ctx():
    # torch ships its own HIP/HSA runtime; tests that hand torch device tensors to the library need it
    # initialised before the library's context opens the device.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    import openfly_amd as fa
    return fa.Context(0)
