import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def pm():
    import parmoon2022_amd
    return parmoon2022_amd
