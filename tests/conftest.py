import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmisohip.so")
    # torch's HIP runtime comes up first, whatever test runs first: the C++
    # layer's test library links the system runtime, and a process whose
    # first HIP user is that library finds torch.cuda unavailable after it
    gpu_available()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    if not gpu_available():
        pytest.skip("no GPU")
    from miso_amd import mhd
    c = mhd.ExecContext(0)
    yield c
    c.close()
