import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libzeknox.so")


@pytest.fixture(scope="session")
def zk():
    """The HIP library binding; GPU tests only."""
    import zeknox_amd
    if zeknox_amd.get_number_of_gpus_rs() < 1:
        pytest.fail("no HIP device visible to libzeknox")
    return zeknox_amd
