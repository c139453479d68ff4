import os
import sys

import pytest

try:  # one HIP runtime per process: torch's, loaded before libspai_mi355x
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    import dcu_amd as D

    c = D.Context(0)
    yield c
    c.close()


def reference_available() -> bool:
    return os.path.isdir(os.path.join(REFERENCE, "matrices"))
