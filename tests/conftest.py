import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblst_amd.so")


@pytest.fixture(scope="session")
def geom_dir():
    from tracklooper_amd import synth
    return synth.geometry_dir()


@pytest.fixture(scope="session")
def oracle_geometry(geom_dir):
    from oracle import orc
    orc.build()
    return orc.Geometry(geom_dir)
