import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

FIXTURES = Path(__file__).parent / "golden" / "fixtures.npz"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: long-running CPU oracle test")


@pytest.fixture(scope="session")
def fx():
    with np.load(FIXTURES, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def mol(fx, key):
    return fx[f"{key}_numbers"], fx[f"{key}_positions"]
