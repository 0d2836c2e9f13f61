import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built libskinny.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The product library on cuda:0; fails (never skips) when it is missing."""
    import torch
    import skinny_matmul_amd as sk
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    assert sk.available(), "libskinny.so is not built"
    sk.lib()
    return sk
