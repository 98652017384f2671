import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def ctx():
    from trl_amd import _lib

    c = _lib.Context(int(os.environ.get("TRL_HIP_DEVICE", "0")))
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
