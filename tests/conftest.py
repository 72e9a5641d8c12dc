import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device)")


@pytest.fixture(scope="session")
def adi():
    from phd_amd import chemdata as cd
    return cd.Adim()


@pytest.fixture(scope="session")
def thd(adi):
    from phd_amd import chemdata as cd
    return cd.build_thd(adi)


@pytest.fixture(scope="session")
def mech():
    from phd_amd import chemdata as cd
    return cd.read_mechanism()


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()
