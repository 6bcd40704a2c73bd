import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnabo_hip.so)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc

    return orc.default()


@pytest.fixture(scope="session")
def nabo():
    """The HIP library; GPU tests fail (not skip) when it or the GPU is missing."""
    import libnabo_amd

    libnabo_amd.load_library()
    assert libnabo_amd.device_count() > 0, "no HIP device visible"
    return libnabo_amd


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
