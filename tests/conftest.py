"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built
libsep.so; everything else runs on the CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libsep.so")


@pytest.fixture(scope="session")
def product():
    from seplib_amd.sim import ProductAPI
    return ProductAPI()


@pytest.fixture(scope="session")
def oracle():
    from oracle.checkers import OracleAPI
    return OracleAPI()


@pytest.fixture(scope="session")
def reference():
    from oracle.checkers import ReferenceAPI
    if not ReferenceAPI.available():
        pytest.skip("compiled reference not built (oracle/_ref)")
    return ReferenceAPI()
