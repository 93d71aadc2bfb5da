import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built native library")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def gpu(pkg):
    """The GPU tests fail (not skip) when no device is visible."""
    import torch
    n = pkg.device_count()
    assert n >= 1, "no HIP device visible"
    assert torch.cuda.is_available()
    return torch.device("cuda:0")
