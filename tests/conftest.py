"""Shared test set-up: markers, import paths, the oracle and GPU fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import py_oracle
    py_oracle.load()
    return py_oracle


@pytest.fixture(scope="session")
def ctx():
    import shinobi_amd as sa
    if sa.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    c = sa.Context(0)
    yield c
    c.close()
