"""Shared test setup: import paths, the `gpu` marker, common helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "sphtogrid.jl_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def max_rel(a, b, floor_frac=1e-3):
    """max |a-b| / |b| over pixels with |b| above floor_frac * max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b) > floor_frac * np.abs(b).max() if b.size else np.zeros(0, bool)
    if not np.any(m):
        return 0.0
    return float((np.abs(a - b)[m] / np.abs(b)[m]).max())


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def s2g():
    import sph2grid
    return sph2grid
