import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dem-engine-sbel_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU runs of the oracle")


@pytest.fixture(scope="session")
def lib():
    import deme_amd

    return deme_amd.load()


@pytest.fixture(scope="session")
def ora_lib():
    import pyoracle

    return pyoracle.load()
