import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm GPU) and the built HIP library')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session')
def gpu():
    """Fails loudly (rather than skipping) when a GPU test runs without a
    GPU or without the HIP library: GPU tests must exercise native code."""
    import torch
    assert torch.cuda.is_available(), 'GPU test run without a ROCm GPU'
    from so3g_amd import _native
    _native.load()
    _native.init()
    return torch.device('cuda')
