"""Shared test setup: import paths, the `gpu` marker and common fixtures."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "spai-gmres-ir_amd"))
sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    import bspai_hip

    return bspai_hip.Context(0)


@pytest.fixture(scope="session")
def c1_golden():
    return dict(np.load(os.path.join(GOLDEN, "c1_reference.npz")))


@pytest.fixture(scope="session")
def band_golden():
    return dict(np.load(os.path.join(GOLDEN, "band_reference.npz")))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)
