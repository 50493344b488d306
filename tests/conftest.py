import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "vulkan-light-bakery_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvlb.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def ctx():
    import vlb

    c = vlb.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    """Per-probe, per-channel relative L2 error over coefficients: ||a-b|| / ||b||."""
    import numpy as np

    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = np.sqrt(((a - b) ** 2).sum(axis=-2))
    den = np.sqrt((b ** 2).sum(axis=-2))
    return num / np.maximum(den, 1e-30)
