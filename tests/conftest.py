import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_phantom():
    from mrtrix3_amd import phantom
    return phantom.single_shell_small()


@pytest.fixture(scope="session")
def small_oracle(small_phantom):
    from oracle import orc
    return orc.CsdOracle(small_phantom.grad, small_phantom.response)
