import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def synth():
    import __graft_entry__ as ge
    return ge._load_file("aldic_synth", os.path.join(ge.PKG_DIR, "synth.py"))


@pytest.fixture(scope="session")
def mesh():
    import __graft_entry__ as ge
    return ge._load_file("aldic_mesh", os.path.join(ge.PKG_DIR, "mesh.py"))
