"""Adjoint of the HipFull field map and its torch.autograd caller (needs one MI355X).

The adjoint is the exact transpose of the plan's FMM operator, so:
  * <y, A x> = <A^T y, x> holds to FP64 rounding on an FP64 plan;
  * A^T y equals the oracle's FMM with sources and targets exchanged on the
    same root cube and depth (the transposed operator chain is that FMM), to
    FP64 rounding, and the FP64 direct sum with exchanged roles to that FMM's
    own truncation error;
  * torch gradients of a linear loss through dipole_field equal A^T w and
    central differences in FP64 along random directions.
"""

import numpy as np
import pytest

from conftest import rel_l2

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import cdfmm_hip

    assert cdfmm_hip.hip_full_available(), "no gfx950 device is visible to the HIP runtime"
    return cdfmm_hip


def make(hip, src, tgt, p, depth, prec, ids=None, mode=None, root=None):
    o = hip.UniformFmmOptions()
    o.expansion_order = p
    o.tree.max_level = depth
    o.backend = hip.ExecutionBackend.HIP_FULL
    o.precision = hip.StaticPrecision.FLOAT32 if prec == 32 else hip.StaticPrecision.FLOAT64
    if ids is not None:
        o.fixed_target_source_indices = [int(v) for v in ids]
    if mode is not None:
        o.hip_m2l_mode = mode
    if root is not None:
        o.tree.root_centre = hip.Vec3(*[float(v) for v in root[:3]])
        o.tree.root_half_width = float(root[3])
    return hip.UniformFmm(src, tgt, o)


def announce(*what):
    print("case:", *what, flush=True)


def slab_geometry(rng, ns, side):
    """Sources in a thin slab, targets on a jittered grid above it (no identities)."""
    src = rng.uniform(0, 1, size=(ns, 3)) * np.array([1.0, 1.0, 0.1])
    g = (np.arange(side) + 0.5) / side
    gx, gy = np.meshgrid(g, g, indexing="ij")
    tgt = np.stack([gx.ravel(), gy.ravel(), np.full(side * side, 0.2)], axis=1)
    tgt[:, :2] += rng.uniform(-0.1, 0.1, size=(side * side, 2)) / side
    return src, tgt


def root_of(f):
    c = f.tree.root_centre
    return np.array([c.x, c.y, c.z, f.tree.root_half_width])


@pytest.mark.parametrize("case", ["identity", "separate"])
def test_adjoint_dot_product_fp64(hip, case):
    rng = np.random.default_rng(700 + len(case))
    if case == "identity":
        src = rng.uniform(0, 1, size=(6000, 3))
        tgt, ids = src, np.arange(6000)
    else:
        src, tgt = rng.uniform(0, 1, size=(5000, 3)), rng.uniform(0.2, 1.2, size=(3000, 3))
        ids = None
    K = 3
    announce("adjoint dot product", case)
    f = make(hip, src, tgt, 6, 3, 64, ids=ids)
    x = rng.normal(size=(K, len(src), 3))
    y = rng.normal(size=(K, len(tgt), 3))
    ax = f.evaluate_batch(x)
    aty = f.evaluate_adjoint_batch(y)
    for k in range(K):
        lhs, rhs = np.vdot(y[k], ax[k]), np.vdot(aty[k], x[k])
        scale = np.linalg.norm(y[k]) * np.linalg.norm(ax[k])
        print(f"  state {k}: <y,Ax> {lhs:.15e} <A^T y,x> {rhs:.15e} rel {abs(lhs - rhs) / scale:.2e}", flush=True)
        assert abs(lhs - rhs) <= 1e-13 * scale


@pytest.mark.parametrize("prec,bar", [(64, 1e-12), (32, 1e-5)])
def test_adjoint_equals_role_exchanged_fmm(hip, prec, bar):
    # Same root cube and depth with the roles exchanged: the oracle's FMM of
    # the transposed problem is the transposed operator chain.
    rng = np.random.default_rng(710 + prec)
    src, tgt = rng.uniform(0, 1, size=(4000, 3)), rng.uniform(0.1, 0.9, size=(2500, 3))
    announce("adjoint vs exchanged oracle", prec)
    f = make(hip, src, tgt, 4, 3, prec)
    root = root_of(f)
    y = rng.normal(size=(2, len(tgt), 3))
    aty = f.evaluate_adjoint_batch(y)
    ref = O.CpuFmm(tgt, src, 4, 3, 64, root_centre=root[:3], root_half_width=root[3])
    for k in range(2):
        e = rel_l2(aty[k], ref.evaluate(y[k])["H"])
        print(f"  state {k}: rel L2 vs exchanged FP64 FMM {e:.3e}", flush=True)
        assert e <= bar


@pytest.mark.parametrize("mode", ["MFMA_BF16X3", "MFMA_FP16X3"])
def test_batched_adjoint_on_matrix_cores(hip, mode):
    # K = 32: the adjoint's near field and M2L run on the matrix cores too.
    rng = np.random.default_rng(715)
    n = 12000
    x = rng.uniform(0, 1, size=(n, 3))
    announce("batched adjoint", mode)
    f = make(hip, x, x, 6, 3, 32, ids=np.arange(n), mode=getattr(hip.HipM2LMode, mode))
    y = rng.normal(size=(32, n, 3))
    aty = f.evaluate_adjoint_batch(y)
    root = root_of(f)
    ref = O.CpuFmm(x, x, 6, 3, 64, root_centre=root[:3], root_half_width=root[3], fixed_ids=np.arange(n))
    for k in (0, 31):
        e = rel_l2(aty[k], ref.evaluate(y[k])["H"])
        print(f"  state {k}: rel L2 vs FP64 FMM {e:.3e}", flush=True)
        assert e <= 1e-5


def test_batched_adjoint_of_separate_points_on_matrix_cores(hip):
    # Sources and targets apart at K = 32 on an FP32 plan: the adjoint's near
    # field walks the block lists of the role-exchanged view (the blocks of
    # the target leaves around each source leaf).
    rng = np.random.default_rng(716)
    src, tgt = rng.uniform(0, 1, size=(5000, 3)), rng.uniform(0.05, 0.95, size=(3500, 3))
    announce("batched adjoint, separate points")
    f = make(hip, src, tgt, 5, 3, 32, mode=hip.HipM2LMode.MFMA_BF16X3)
    y = rng.normal(size=(32, len(tgt), 3))
    aty = f.evaluate_adjoint_batch(y)
    assert aty.shape == (32, len(src), 3)
    root = root_of(f)
    ref = O.CpuFmm(tgt, src, 5, 3, 64, root_centre=root[:3], root_half_width=root[3])
    for k in (0, 31):
        e = rel_l2(aty[k], ref.evaluate(y[k])["H"])
        print(f"  state {k}: rel L2 vs exchanged FP64 FMM {e:.3e}", flush=True)
        assert e <= 1e-5


def test_adjoint_identity_map_is_transposed(hip):
    # Targets = sources with a shuffled identity map: the adjoint skips the
    # pair (source s, target t) exactly where the forward skips (t, s).
    rng = np.random.default_rng(720)
    n = 3000
    x = rng.uniform(0, 1, size=(n, 3))
    tgt_perm = rng.permutation(n)
    tgt = x[tgt_perm]
    ids = tgt_perm  # target i is source tgt_perm[i]
    announce("adjoint shuffled identity")
    f = make(hip, x, tgt, 4, 2, 64, ids=ids)
    y = rng.normal(size=(1, n, 3))
    aty = f.evaluate_adjoint_batch(y)[0]
    inv = np.empty(n, dtype=np.int64)
    inv[tgt_perm] = np.arange(n)  # source s is target inv[s]
    ref = O.CpuFmm(tgt, x, 4, 2, 64, root_centre=root_of(f)[:3], root_half_width=root_of(f)[3], fixed_ids=inv)
    e = rel_l2(aty, ref.evaluate(y[0])["H"])
    print(f"  rel L2 vs exchanged FP64 FMM with the inverse map {e:.3e}", flush=True)
    assert e <= 1e-12
    # Two targets on one source: no transposed identity map.
    g = make(hip, x[:10], x[[0, 0, 1]], 2, 1, 64, ids=[0, 0, 1])
    with pytest.raises(ValueError, match="two targets with one source"):
        g.evaluate_adjoint_batch(np.ones((1, 3, 3)))


def test_adjoint_slab_far_field_fp32(hip):
    # Sources in a thin slab, targets on a grid above it: the near field is
    # empty at depth 5 and the gradient is all far field.
    rng = np.random.default_rng(730)
    src, tgt = slab_geometry(rng, 40000, 64)
    announce("adjoint slab far field")
    f32 = make(hip, src, tgt, 6, 5, 32)
    f64 = make(hip, src, tgt, 6, 5, 64)
    y = rng.normal(size=(2, len(tgt), 3))
    g32, g64 = f32.evaluate_adjoint_batch(y), f64.evaluate_adjoint_batch(y)
    for k in range(2):
        e = rel_l2(g32[k], g64[k])
        print(f"  state {k}: FP32 vs FP64 adjoint {e:.3e}", flush=True)
        assert e <= 1e-5
    root = root_of(f64)
    ref = O.CpuFmm(tgt, src, 6, 5, 64, root_centre=root[:3], root_half_width=root[3]).evaluate(y[0])["H"]
    e = rel_l2(g64[0], ref)
    print(f"  FP64 adjoint vs exchanged FP64 FMM {e:.3e}", flush=True)
    assert e <= 1e-12
    # All-far-field geometry: p = 6 truncation is ~5e-3 here (the oracle's own
    # FMM shows the same against the direct sum), so the direct sum checks only
    # that the adjoint carries that error and no more.
    direct = O.direct(src[:2000], tgt, y[0])["H"]
    e_fmm, e_ref = rel_l2(g64[0][:2000], direct), rel_l2(ref[:2000], direct)
    print(f"  vs exchanged direct sum (2000 sources): adjoint {e_fmm:.3e} oracle FMM {e_ref:.3e}", flush=True)
    assert e_fmm <= 1.01 * e_ref


@pytest.mark.parametrize("prec", [32, 64])
def test_torch_autograd_through_dipole_field(hip, prec):
    import torch

    rng = np.random.default_rng(740 + prec)
    n = 5000
    x = rng.uniform(0, 1, size=(n, 3))
    announce("torch autograd", prec)
    f = make(hip, x, x, 6, 3, prec, ids=np.arange(n))
    dt = torch.float32 if prec == 32 else torch.float64
    m = torch.tensor(rng.normal(size=(2, n, 3)), dtype=dt, device="cuda", requires_grad=True)
    w = torch.tensor(rng.normal(size=(2, n, 3)), dtype=dt, device="cuda")
    h = hip.dipole_field(f, m)
    loss = (w * h).sum()
    loss.backward()
    grad = m.grad.detach().cpu().numpy()
    want = f.evaluate_adjoint_batch(w.cpu().numpy().astype(np.float64))
    for k in range(2):
        assert rel_l2(grad[k], want[k]) <= (1e-6 if prec == 32 else 1e-14)
    # Central differences along random directions (exact for a linear map up
    # to rounding), in the plan precision.
    if prec == 64:
        for d in range(8):
            v = torch.tensor(rng.normal(size=(2, n, 3)), dtype=dt, device="cuda")
            with torch.no_grad():
                eps = 1e-3
                lp = (w * hip.dipole_field(f, m + eps * v)).sum()
                lm = (w * hip.dipole_field(f, m - eps * v)).sum()
                fd = float((lp - lm) / (2 * eps))
            an = float((m.grad * v).sum())
            assert abs(fd - an) <= 1e-10 * max(abs(an), 1.0), (d, fd, an)
    # Second order: the backward is itself differentiable.  g = A^T w as a
    # function of w, so d<g, u>/dw = A u, the forward map again.
    w2 = w.detach().clone().requires_grad_(True)
    h2 = hip.dipole_field(f, m)
    (g,) = torch.autograd.grad((w2 * h2).sum(), m, create_graph=True)
    u = torch.tensor(rng.normal(size=(2, n, 3)), dtype=dt, device="cuda")
    (gw,) = torch.autograd.grad((g * u).sum(), w2)
    au = hip.dipole_field(f, u).detach().cpu().numpy()
    for k in range(2):
        assert rel_l2(gw.detach().cpu().numpy()[k], au[k]) <= (1e-6 if prec == 32 else 1e-14)
