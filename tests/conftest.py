import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def pkg():
    import helpers
    return helpers.load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import helpers
    return helpers.load_oracle()
