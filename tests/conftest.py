import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

XSEC_FILE = os.path.join(ROOT, "boltzmann_amd", "data", "eAr_synthetic_3proc.lxcat")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libbz.so")


@pytest.fixture(scope="session")
def xsec_file():
    return XSEC_FILE
