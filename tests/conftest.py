import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long-running case")
    os.environ.setdefault("EXOBLAS_QUIET", "1")


@pytest.fixture
def rng():
    import numpy as np
    return np.random.default_rng(0)
