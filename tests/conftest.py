"""Shared test set-up.  Tests marked ``gpu`` need an MI355X; everything else
runs on the CPU (the oracle, fixtures, host logic, library exports)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


@pytest.fixture(scope="session")
def tiny_scan():
    from phasecon_amd.synth import make_scan
    return make_scan("tiny3d", seed=0)


@pytest.fixture(scope="session")
def small4d_scan():
    from phasecon_amd.synth import make_scan
    return make_scan("ch4d_small", seed=3)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor, 1e-300))
