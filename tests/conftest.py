import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def literals():
    with open(os.path.join(GOLDEN, "ref_literals.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def refdata():
    def read(name):
        with open(os.path.join(GOLDEN, "reference_data", name)) as fh:
            return fh.read()
    return read


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle
