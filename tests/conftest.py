import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# PyTorch-ROCm ships its own HIP runtime next to the one the engine links
# (/opt/rocm).  In one process torch's runtime has to reach the device
# first: initialised after the engine's, it finds no device.  The GPU tests
# that hand torch tensors to the engine therefore rely on this call, made
# before any test creates a context (on a host without a GPU it only
# returns False).
try:
    import torch
    torch.cuda.is_available()
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs through the HIP engine)")
    config.addinivalue_line(
        "markers", "slow: minutes of CPU work")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def synth():
    from cvt_amd import synth as S
    return S
