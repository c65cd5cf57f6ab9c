import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

import __graft_entry__ as entry  # noqa: E402

# registers the package directory as `geothermalwells_jl_amd`
entry.load_package()
entry.load_oracle()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def gw():
    return entry.load_package()


@pytest.fixture(scope="session")
def O():
    return entry.load_oracle()
