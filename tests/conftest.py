import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def dev():
    import torch
    import bale_old_amd as B
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    B.check(B.lib().bale_hip_init(0))
    return torch.device("cuda:0")
