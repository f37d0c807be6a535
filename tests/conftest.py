import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sdf-slam_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def sdfslam():
    import sdfslam as mod
    import oracle_ctypes
    oracle_ctypes.configure(mod)
    return mod


@pytest.fixture(scope="session")
def oracle(sdfslam):
    import oracle_ctypes
    return oracle_ctypes
