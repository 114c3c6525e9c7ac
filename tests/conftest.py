import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    """The FP64 CPU oracle (built from oracle/ref_cpu.c on first use)."""
    from oracle import ref_cpu

    ref_cpu.lib()
    return ref_cpu


@pytest.fixture(scope="session")
def bbx():
    """The gfx950 library; GPU tests fail loudly if it is not built."""
    from bubbles_amd import abi

    abi.load()
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return abi
