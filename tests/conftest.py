import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def fixture20k(oracle):
    """Data/UTest/unitCubeRef20kDouble.bfma: 20k particles with FP64 direct results."""
    return oracle.read_bfma(os.path.join(GOLDEN, "unitCubeRef20kDouble.bfma"))


@pytest.fixture(scope="session")
def fixture5k(oracle):
    """Data/UTest/DirectDouble.bfma: 5k particles (q=0.1) with FP64 direct results."""
    return oracle.read_bfma(os.path.join(GOLDEN, "DirectDouble.bfma"))
