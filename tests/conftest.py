import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(ROOT, "oracle")
for _p in (TESTS, ORACLE):
    if _p not in sys.path:
        sys.path.insert(0, _p)

DATA = os.path.join(ROOT, "tests", "golden", "racon_data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def data_dir():
    return DATA
