import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def lri():
    """The host package (directory ``lowrankintegrators.jl_amd``), imported by path."""
    import __graft_entry__

    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def native(lri):
    """The loaded C-ABI library; a GPU test fails loudly if it is missing."""
    return lri.native.lib()
