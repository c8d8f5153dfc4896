"""Shared pytest set-up: the `gpu` marker, import paths and fixtures.

`-m "not gpu"` runs the oracle pins and host-side checks on any machine;
`-m gpu` runs the parity tests that call libscuffx.so on a GPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "scuff-em_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (calls libscuffx.so kernels)")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def sx():
    import scuffx
    return scuffx


def rel_fro(A, B):
    """Relative Frobenius distance ||A-B|| / ||B||."""
    return float(np.linalg.norm(np.asarray(A) - np.asarray(B)) / np.linalg.norm(np.asarray(B)))
