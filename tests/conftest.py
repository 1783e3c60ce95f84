import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; GPU tests then run through the C ABI."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from gsparc_amd import sptc
    sptc.lib()
    return torch.device("cuda:0")
