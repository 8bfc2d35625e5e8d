import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def panda():
    from rtcollisiondetection_amd import robots
    return robots.load_panda()


@pytest.fixture(scope="session")
def orc_panda(panda):
    from oracle import oracle as O
    return O.Robot(panda)
