import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

CASES = os.path.join(HERE, "golden", "cases")
GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def case_par(name):
    return os.path.join(CASES, name, "Par.inp")


@pytest.fixture(scope="session")
def oracle_api():
    import oracle
    return oracle.load()


@pytest.fixture(scope="session")
def oracle_ulp_api():
    """the oracle with every libm result moved by one ulp (conditioning probe)"""
    import oracle
    return oracle.load(ulp_perturbed=True)
