"""Field gradient and position gradients of the HipFull path (needs one
MI355X).  No counterpart in the reference; the bars:

* near field alone (a depth-1 tree, where every pair is a neighbour pair):
  the FP64 closed-form direct gradient, itself pinned against differences of
  the oracle's direct field, to 1e-12 (FP64) and 1e-5 (FP32);
* the full FMM: central differences of the plan's own FP64 field under
  target displacements that keep every target in its leaf, to 1e-6; the FP64
  direct gradient to 1e-6 (p = 8, where the near field dominates the norm);
* source positions: the directional derivative of L = sum g . H under source
  displacements that keep every source in its leaf;
* the adjoint field gradient: the forward field gradient of the plan with
  sources and targets exchanged, to 1e-12.
"""

import numpy as np
import pytest

from conftest import direct_field_gradient, gradient_matrix, rel_l2

pytestmark = pytest.mark.gpu

UNIT_BOX = ((0.5, 0.5, 0.5), 0.5)


@pytest.fixture(scope="module")
def hip():
    import cdfmm_hip

    assert cdfmm_hip.hip_full_available(), "no gfx950 device is visible to the HIP runtime"
    return cdfmm_hip


def make(hip, src, tgt, p, depth, prec, ids=None):
    o = hip.UniformFmmOptions()
    o.expansion_order = p
    o.tree.max_level = depth
    o.tree.root_centre = hip.Vec3(*UNIT_BOX[0])
    o.tree.root_half_width = UNIT_BOX[1]
    o.backend = hip.ExecutionBackend.HIP_FULL
    o.precision = hip.StaticPrecision.FLOAT32 if prec == 32 else hip.StaticPrecision.FLOAT64
    if ids is not None:
        o.fixed_target_source_indices = [int(v) for v in ids]
    return hip.UniformFmm(src, tgt, o)


def announce(*what):
    print("case:", *what, flush=True)


def inside_leaves(x, depth, margin):
    """Mask of points at least `margin` away from every leaf face of the unit box."""
    w = 1.0 / (1 << depth)
    f = np.mod(x, w)
    return np.all((f > margin) & (f < w - margin), axis=1)


@pytest.mark.parametrize("prec,K", [(64, 3), (32, 3), (32, 32)])
def test_near_field_gradient_matches_the_direct_gradient(hip, prec, K):
    # K = 32 on an FP32 plan runs the near-field gradient on bf16
    # three-piece operands; K = 3 the vector kernel.
    rng = np.random.default_rng(1201)
    n = 2500
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    announce("near-field gradient", prec, K)
    f = make(hip, x, x, 4, 1, prec, ids=np.arange(n))
    G = f.evaluate_field_gradient_batch(ms)
    assert G.shape == (K, n, 6) and G.dtype == (np.float32 if prec == 32 else np.float64)
    for k in sorted({0, 1, K - 1}):
        want = direct_field_gradient(x, x, ms[k], np.arange(n))
        e = rel_l2(G[k], want)
        print(f"  state {k}: rel L2 {e:.3e}", flush=True)
        assert e <= (1e-12 if prec == 64 else 1e-5)


def test_fmm_gradient_is_the_derivative_of_the_fmm_field(hip):
    # Targets apart from the sources, each well inside its leaf: moving all of
    # them at once by +-h along an axis leaves the tree unchanged, and the
    # field at one target does not depend on the others.
    rng = np.random.default_rng(1202)
    depth, p, K = 3, 6, 2
    src = rng.uniform(0, 1, size=(4000, 3))
    tgt = rng.uniform(0, 1, size=(3000, 3))
    tgt = tgt[inside_leaves(tgt, depth, 1e-3)][:400]
    ms = rng.normal(size=(K, len(src), 3))
    announce("FMM gradient vs differences", len(tgt))
    G = gradient_matrix(make(hip, src, tgt, p, depth, 64).evaluate_field_gradient_batch(ms))
    h = 3e-7
    fd = np.zeros_like(G)
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        hp = make(hip, src, tgt + e, p, depth, 64).evaluate_batch(ms)
        hm = make(hip, src, tgt - e, p, depth, 64).evaluate_batch(ms)
        fd[..., j] = (hp - hm) / (2 * h)
    err = rel_l2(G, fd)
    print(f"  rel L2 vs central differences {err:.3e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("prec", [64, 32])
def test_fmm_gradient_against_the_direct_gradient(hip, prec):
    rng = np.random.default_rng(1203)
    n, K, p, depth = 6000, 32, 8, 3
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    announce("FMM gradient vs direct", prec)
    f = make(hip, x, x, p, depth, prec, ids=np.arange(n))
    field_before = f.evaluate_batch(ms[:4])
    G = f.evaluate_field_gradient_batch(ms)
    for k in (0, K - 1):
        want = direct_field_gradient(x, x, ms[k], np.arange(n))
        e = rel_l2(G[k], want)
        tr = np.abs(G[k][:, 0] + G[k][:, 3] + G[k][:, 5]).max() / np.abs(G[k]).max()
        print(f"  state {k}: rel L2 vs direct {e:.3e}, max |trace| / max |G| {tr:.2e}", flush=True)
        assert e <= (1e-6 if prec == 64 else 2e-6) and tr <= (1e-12 if prec == 64 else 1e-6)
    if prec == 32:
        G64 = make(hip, x, x, p, depth, 64, ids=np.arange(n)).evaluate_field_gradient_batch(ms)
        e = rel_l2(G, G64)
        print(f"  FP32 vs FP64 FMM gradient {e:.3e}", flush=True)
        assert e <= 1e-5
    # The field of the same plan is unchanged after the gradient evaluate.
    np.testing.assert_array_equal(f.evaluate_batch(ms[:4]), field_before)


def test_source_and_target_position_gradients(hip):
    import torch

    import cdfmm_hip

    rng = np.random.default_rng(1204)
    depth, p, K = 3, 6, 2
    src = rng.uniform(0, 1, size=(3000, 3))
    tgt = rng.uniform(0, 1, size=(2500, 3))
    ms = rng.normal(size=(K, len(src), 3))
    gs = rng.normal(size=(K, len(tgt), 3))
    announce("position gradients")
    f = make(hip, src, tgt, p, depth, 64)
    d_t, d_s = cdfmm_hip.position_gradients(f, torch.tensor(ms, device="cuda"), torch.tensor(gs, device="cuda"))
    d_t, d_s = d_t.cpu().numpy(), d_s.cpu().numpy()
    assert d_t.shape == (len(tgt), 3) and d_s.shape == (len(src), 3)

    def loss(s, t):
        return float(np.sum(gs * make(hip, s, t, p, depth, 64).evaluate_batch(ms)))

    h = 1e-7
    for name, pts, grad in (("sources", src, d_s), ("targets", tgt, d_t)):
        delta = rng.normal(size=pts.shape) * inside_leaves(pts, depth, 1e-3)[:, None]
        if name == "sources":
            fd = (loss(src + h * delta, tgt) - loss(src - h * delta, tgt)) / (2 * h)
        else:
            fd = (loss(src, tgt + h * delta) - loss(src, tgt - h * delta)) / (2 * h)
        an = float(np.sum(grad * delta))
        print(f"  {name}: directional derivative {an:.9e}, differences {fd:.9e}", flush=True)
        assert abs(an - fd) <= 1e-6 * abs(an)


@pytest.mark.parametrize("prec", [64, 32])
def test_adjoint_field_gradient_is_the_exchanged_plans_gradient(hip, prec):
    rng = np.random.default_rng(1205)
    src = rng.uniform(0, 1, size=(3000, 3))
    tgt = rng.uniform(0, 1, size=(2000, 3))
    K = 32 if prec == 32 else 3
    gs = rng.normal(size=(K, len(tgt), 3))
    announce("adjoint field gradient", prec)
    fwd = make(hip, src, tgt, 6, 3, prec)
    Ga = fwd.evaluate_field_gradient_adjoint_batch(gs)
    assert Ga.shape == (K, len(src), 6)
    Gx = make(hip, tgt, src, 6, 3, 64).evaluate_field_gradient_batch(gs)
    e = rel_l2(Ga, Gx)
    print(f"  rel L2 vs exchanged FP64 plan {e:.3e}", flush=True)
    assert e <= (1e-12 if prec == 64 else 1e-5)


def test_device_field_gradient_matches_host(hip):
    import torch

    import cdfmm_hip

    rng = np.random.default_rng(1206)
    n, K = 8000, 32
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    announce("device field gradient")
    f = make(hip, x, x, 6, 3, 32, ids=np.arange(n))
    host = f.evaluate_field_gradient_batch(ms)
    dev = cdfmm_hip.field_gradient(f, torch.tensor(ms, device="cuda")).cpu().numpy()
    e = rel_l2(dev, host)
    print(f"  device vs host {e:.3e}", flush=True)
    assert e <= 1e-6
    with pytest.raises(ValueError):
        cdfmm_hip.field_gradient(f, torch.zeros((K, n + 1, 3), device="cuda"))


@pytest.mark.parametrize("prec,K", [(64, 2), (32, 64)])
def test_shared_point_gradients_with_the_identity_map(hip, prec, K):
    # Targets = sources with the identity map: moving a point moves a source
    # and a target, so dL/dx = dL/dx_t + dL/dx_s; against central differences
    # of L along a direction that keeps every point in its leaf (FP64), and
    # against the FP64 gradients (FP32 at K = 64).
    import torch

    import cdfmm_hip

    rng = np.random.default_rng(1207)
    n, depth, p = 4000, 3, 6
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    gs = rng.normal(size=(K, n, 3))
    announce("shared-point gradients", prec, K)

    def grads(prec_):
        f = make(hip, x, x, p, depth, prec_, ids=np.arange(n))
        d_t, d_s = cdfmm_hip.position_gradients(f, torch.tensor(ms, device="cuda"), torch.tensor(gs, device="cuda"))
        return (d_t + d_s).cpu().numpy().astype(np.float64)

    total = grads(prec)
    if prec == 32:
        e = rel_l2(total, grads(64))
        print(f"  FP32 vs FP64 shared-point gradient {e:.3e}", flush=True)
        assert e <= 1e-5
        return

    def loss(pts):
        return float(np.sum(gs * make(hip, pts, pts, p, depth, 64, ids=np.arange(n)).evaluate_batch(ms)))

    h = 1e-7
    delta = rng.normal(size=x.shape) * inside_leaves(x, depth, 1e-3)[:, None]
    fd = (loss(x + h * delta) - loss(x - h * delta)) / (2 * h)
    an = float(np.sum(total * delta))
    print(f"  directional derivative {an:.9e}, differences {fd:.9e}", flush=True)
    assert abs(an - fd) <= 1e-6 * abs(an)
