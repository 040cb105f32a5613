import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("surfatt-iso_amd", "oracle", os.path.join("tests", "golden")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_solvers.npz"))


@pytest.fixture(scope="session")
def ctx():
    """A satt context on GPU 0.  Fails (does not skip) when the HIP library
    or the GPU is missing: GPU tests must exercise the native path."""
    import satt
    c = satt.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    """max |a-b| / max |b|"""
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
