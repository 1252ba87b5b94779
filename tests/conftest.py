import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long CPU-side oracle run")


@pytest.fixture(scope="session")
def mtr():
    return importlib.import_module("multi-tilt-reconstuction_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle  # noqa: WPS433 (the CPU checker)

    return oracle


@pytest.fixture(scope="session")
def ctx(mtr):
    c = mtr.Context(0)
    yield c
    c.close()
