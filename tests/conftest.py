import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libstaqdic.so)")


@pytest.fixture(scope="session")
def ctx():
    from staq_dic_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def plain_cfg():
    """512^2 affine pair on the uniform 900-node grid (seed 1001)."""
    from staq_dic_amd import synth
    return synth.config_plain()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle
