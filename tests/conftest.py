import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip_lib():
    """The gfx950 library (built in-tree by __graft_entry__.build())."""
    from isaac_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from isaac_amd import build

        build.build()
    return _lib.lib()


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu(hip_lib):
    if not gpu_available():
        pytest.fail("GPU test collected without a visible GPU")
    import torch

    torch.cuda.init()
    return torch.device("cuda", 0)
