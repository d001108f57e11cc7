import glob
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
    yield


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def golden_names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def force_tol(scale, rel=1e-10):
    """Per-particle tolerance: rel times the sum of |f_ij| over the particle's
    contacts (the value itself is ill-conditioned at force balance), with a
    floor for particles without contacts."""
    return rel * np.maximum(scale, 1e-300) + 1e-300


def wrap_diff(a, b, L):
    """Difference of coordinates modulo the box length."""
    d = a - b
    return d - L * np.round(d / L)
