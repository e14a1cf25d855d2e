import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhipotf.so)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch


def gpu_vec(torch, a):
    import numpy as np
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def host(t):
    return t.detach().cpu().numpy()
