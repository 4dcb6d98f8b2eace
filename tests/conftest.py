import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU and the built libcrocgpu.so')
    config.addinivalue_line('markers', 'slow: full-size workload parity (minutes of CPU oracle time)')


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('GPU test collected on a machine without a ROCm GPU')
    return torch.device('cuda', 0)
