import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libvv1d.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch sees no GPU")
    return torch.device("cuda")


@pytest.fixture
def progress(capsys):
    """Print a progress line of a long test as it happens, past pytest's
    output capture (so a run that watches its output sees it moving)."""
    def say(*args):
        with capsys.disabled():
            print(*args, flush=True)
    return say
