import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "spid-join_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def sj():
    import spidjoin

    spidjoin.lib()
    return spidjoin


@pytest.fixture(scope="session")
def gpu(sj):
    sj.init(0)
    return sj


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
