import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle32():
    from oracle import Oracle
    return Oracle(32)


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is present; importing torch is not needed for the
    C ABI, but torch.cuda is the simplest probe available on both hosts."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
