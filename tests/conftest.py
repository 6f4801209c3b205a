import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def scale_err(g, c):
    """Max scale-normalised error |g - c| / max|c| (the reference GPU gates,
    step_test.cpp:103-108, force_test.cpp:79,99)."""
    g = np.asarray(g, float)
    c = np.asarray(c, float)
    sc = float(np.max(np.abs(c))) if c.size else 0.0
    d = float(np.max(np.abs(g - c))) if c.size else 0.0
    return d / sc if sc > 0 else d


def rms_scale_err(g, c):
    g = np.asarray(g, float)
    c = np.asarray(c, float)
    sc = float(np.max(np.abs(c)))
    return float(np.sqrt(np.mean((g - c) ** 2))) / sc if sc > 0 else 0.0


def rel_err(g, c, floor=0.0):
    """Per-element relative error |g - c| / max(|c|, floor), maximum over elements."""
    g = np.asarray(g, float)
    c = np.asarray(c, float)
    return float(np.max(np.abs(g - c) / np.maximum(np.abs(c), floor)))


@pytest.fixture(scope="session")
def gpu():
    """The dsph binding; GPU tests fail (not skip) when the HIP library is missing."""
    from deorbit_amd import dsph
    dsph.load_library()
    return dsph
