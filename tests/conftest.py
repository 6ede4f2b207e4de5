import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    from ht_tensor_learning_using_gpu_tensor_cores_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_available():
        pytest.skip("no GPU")
    import torch
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def handle(gpu):
    from ht_tensor_learning_using_gpu_tensor_cores_amd import Handle
    h = Handle(0)
    yield h
    h.close()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
