import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP path through the C ABI")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def mpf():
    import mphyfem_amd
    mphyfem_amd.lib()
    return mphyfem_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(mpf):
    if mpf.device_count() == 0:
        pytest.fail("no HIP device visible: the gpu tests need the MI355X")
    import torch
    assert torch.cuda.is_available()
    return torch.device("cuda:0")
