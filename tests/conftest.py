import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import saber_amd
    c = saber_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx32():
    import saber_amd
    c = saber_amd.Context(0, force_exact32=True)
    yield c
    c.close()
