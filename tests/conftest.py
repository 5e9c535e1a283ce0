"""Shared test setup: import paths for the built HipFull module and the oracle."""

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("dip-fmm_amd", "oracle"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs one MI355X (gfx950) GPU and the built HIP library")


def golden(name: str):
    """Load a committed fixture (plain arrays only; no pickles)."""
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    return oracle
