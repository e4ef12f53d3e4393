import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def lcr():
    """The product package (liblcr.so through ctypes); GPU tests only."""
    import lowcost3dreconstruction_amd as m
    return m


def print_case(**kw):
    """Names the case before a GPU step, so a fault can be traced to it."""
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)
