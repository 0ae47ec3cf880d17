import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    if not os.path.exists(oracle_lib.LIB_PATH):
        oracle_lib.build()
    return oracle_lib.load()


@pytest.fixture(scope="session")
def hostlane():
    """The engine's lane bodies built for the host (tests/host_lane.cpp)."""
    path = os.path.join(ROOT, "tests", "libhostlane.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests")], check=True)
    import hostlane_lib
    return hostlane_lib.load(path)


@pytest.fixture(scope="session")
def engine():
    import egspice_amd
    return egspice_amd.library()
