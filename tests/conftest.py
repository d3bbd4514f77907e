import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(*parts):
    return np.load(os.path.join(GOLDEN, *parts), allow_pickle=False)


@pytest.fixture(scope="session")
def device():
    from biblbm_amd import capi
    capi.call("bl_init", 0)
    return 0
