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


def direct_field_gradient(targets, sources, moments, ids=None):
    """FP64 direct field gradient G_cj = dH_c/dx_j, (n_targets, 6) as
    xx xy xz yy yz zz, of point dipoles with H = (3 u (u.m) - m)/(4 pi r^3):
    G_cj = 3/(4 pi r^4) (delta_cj w + u_c m_j + m_c u_j - 5 u_c u_j w),
    w = u.m; the pair of a target with its own source (ids) is skipped."""
    t = np.asarray(targets, dtype=np.float64)
    s = np.asarray(sources, dtype=np.float64)
    m = np.asarray(moments, dtype=np.float64)
    out = np.zeros((len(t), 6))
    pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
    for i0 in range(0, len(t), 256):
        r = t[i0:i0 + 256, None, :] - s[None, :, :]
        d = np.linalg.norm(r, axis=2)
        if ids is not None:
            own = np.asarray(ids[i0:i0 + 256])
            d[np.arange(len(own)), own] = np.inf
        u = r / d[..., None]
        w = np.einsum("tsc,sc->ts", u, m)
        g = 3.0 / (4.0 * np.pi * d ** 4)
        for e, (c, j) in enumerate(pairs):
            term = (c == j) * w + u[..., c] * m[None, :, j] + m[None, :, c] * u[..., j] - 5.0 * u[..., c] * u[..., j] * w
            out[i0:i0 + 256, e] = np.sum(g * term, axis=1)
    return out


def gradient_matrix(g6):
    """(..., 6) stored entries -> (..., 3, 3) symmetric matrices."""
    g6 = np.asarray(g6)
    idx = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 5]])
    return g6[..., idx]
