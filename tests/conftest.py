import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: test needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def reference_data():
    """Directory of the reference's own test fixtures (read in place)."""
    path = os.path.join(REFERENCE, "testsuite", "python", "data")
    if not os.path.isdir(path):
        pytest.skip("reference test data not present on this machine")
    return path
