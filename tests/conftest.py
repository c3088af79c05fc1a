import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, 'tests')):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device and the native library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def golden():
    import numpy as np

    def load(name):
        path = os.path.join(ROOT, 'tests', 'golden', name + '.npz')
        with np.load(path, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return load
