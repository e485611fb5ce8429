import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line(
        "markers", "slow: oracle runs of more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle import lcode_oracle
    return lcode_oracle


@pytest.fixture(scope="session")
def po():
    from oracle import predictor_oracle
    return predictor_oracle
