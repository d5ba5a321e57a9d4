import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def atg():
    from atg_testlib import load_atg
    return load_atg()


@pytest.fixture(scope="session")
def oracle():
    from atg_testlib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ctx(atg):
    c = atg.Context()
    yield c
    c.close()
