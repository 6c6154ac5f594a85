import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libqtxhip.so")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly (not skip) when selected without a device, so a
    # missing GPU or library is never mistaken for a pass.
    pass
