import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhtgx.so on the device)")


@pytest.fixture(scope="session")
def synth0():
    from htrdr_kmin_amd import atmosphere as A
    s = A.Synth(0, seed=0)
    yield s
    s.close()


@pytest.fixture(scope="session")
def sky0(synth0):
    import oracle_ffi as O
    sky = O.Sky(synth0.scene_ptr)
    yield sky
    sky.close()


@pytest.fixture(scope="session")
def ctx0(synth0):
    from htrdr_kmin_amd import atmosphere as A
    with A.Context(synth0.scene_ptr) as ctx:
        yield ctx
