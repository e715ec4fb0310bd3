import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

import __graft_entry__  # noqa: E402

GOLDEN = REPO / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def cloth():
    """The product package; its shared library must already be built."""
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.library()
    return pyoracle


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
