import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sc25-stz2_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not os.path.exists(pyoracle._LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def stz():
    import stz_amd
    stz_amd.lib()  # raises when libstz_hip.so is not built
    return stz_amd


@pytest.fixture(scope="session")
def ctx(stz):
    c = stz.Context(0)
    yield c
    c.close()
