"""The closed-form field gradient used by the position-gradient tests, pinned
on the CPU against central differences of the oracle's FP64 direct field
(validation.cpp:71-111), which is itself pinned to the reference."""

import numpy as np

from conftest import direct_field_gradient, gradient_matrix, rel_l2

import oracle as O


def test_closed_form_gradient_matches_differences_of_the_direct_field():
    rng = np.random.default_rng(41)
    src = rng.uniform(0, 1, size=(400, 3))
    mom = rng.normal(size=(400, 3))
    tgt = rng.uniform(0.2, 0.8, size=(30, 3))
    G = gradient_matrix(direct_field_gradient(tgt, src, mom))
    h = 1e-5
    fd = np.zeros((30, 3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        hp = O.direct(tgt + e, src, mom)["H"]
        hm = O.direct(tgt - e, src, mom)["H"]
        fd[:, :, j] = (hp - hm) / (2 * h)
    assert rel_l2(G, fd) <= 1e-6
    # Curl-free and divergence-free away from the sources.
    assert np.max(np.abs(np.trace(G, axis1=1, axis2=2))) <= 1e-9 * np.max(np.abs(G))


def test_closed_form_gradient_skips_the_own_source():
    rng = np.random.default_rng(42)
    x = rng.uniform(0, 1, size=(50, 3))
    mom = rng.normal(size=(50, 3))
    ids = np.arange(50)
    full = direct_field_gradient(x, x, mom, ids)
    for t in (0, 17, 49):
        others = np.delete(np.arange(50), t)
        one = direct_field_gradient(x[t:t + 1], x[others], mom[others])
        np.testing.assert_allclose(full[t], one[0], rtol=1e-12, atol=0)
