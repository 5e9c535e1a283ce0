"""Potential output of the HipFull path on the device (needs one MI355X).

The reference's CudaFull path is field-only; CpuStatic returns potentials
(src/uniform_fmm.cpp:1389-1437).  Bars as for the fields: FP32 plans within
rel. L2 1e-5 of CpuStatic, FP64 within 1e-12, against the compiled
reference's 'both' output (tests/golden/ref_potentials.npz) and the oracle.
"""

import numpy as np
import pytest

from conftest import golden, rel_l2

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


BARS = {64: 1e-12, 32: 1e-5}


@pytest.mark.parametrize("case", ["c1", "n4096_p8_d3"])
@pytest.mark.parametrize("prec", [64, 32])
def test_potentials_match_reference_golden(hip, case, prec):
    g = golden("ref_potentials.npz")
    x, m = g[f"{case}/positions"], g[f"{case}/moments"]
    n, p, d = (int(v) for v in g[f"{case}/config"])
    announce("potential golden", case, prec)
    f = make(hip, x, x, p, d, prec, ids=np.arange(n))
    r = f.evaluate(m, output="both")
    e_phi, e_h = rel_l2(r["phi"], g[f"{case}/phi{prec}"]), rel_l2(r["H"], g[f"{case}/H{prec}"])
    print(f"  rel L2 vs reference: phi {e_phi:.3e}  H {e_h:.3e}", flush=True)
    assert e_phi <= BARS[prec] and e_h <= BARS[prec]
    # Potential only: the same potentials, zero fields.
    q = f.evaluate(m, output="potential")
    assert rel_l2(q["phi"], g[f"{case}/phi{prec}"]) <= BARS[prec]
    assert not np.any(q["H"])


def test_separate_targets_potential(hip):
    g = golden("ref_potentials.npz")
    for prec in (64, 32):
        announce("potential separate", prec)
        f = make(hip, g["separate/sources"], g["separate/targets"], 4, 3, prec, root=(0.0, 0.0, 0.0, 1.0))
        r = f.evaluate(g["separate/moments"], output="both")
        assert rel_l2(r["phi"], g[f"separate/phi{prec}"]) <= BARS[prec]
        assert rel_l2(r["H"], g[f"separate/H{prec}"]) <= BARS[prec]


def test_coincident_pair_potential_is_nan(hip):
    g = golden("ref_potentials.npz")
    x, m = g["coincident/positions"], g["coincident/moments"]
    for prec in (64, 32):
        announce("potential coincident", prec)
        phi = make(hip, x, x, 2, 1, prec).evaluate(m, output="potential", target_source_indices=[0, -1, 2, 3])["phi"]
        assert np.isnan(phi[:2]).all() and np.isfinite(phi[2:]).all()
        assert rel_l2(phi[2:], g[f"coincident/phi{prec}"][2:]) <= BARS[prec]


@pytest.mark.parametrize("mode,n", [("VALU", 12000), ("MFMA_BF16X3", 12000), ("MFMA_BF16X3", 8192)])
def test_batched_potentials_on_matrix_cores(hip, mode, n):
    # K = 32 on an FP32 plan: the near-field potential is the fourth row per
    # target of the matrix-core P2P, the far potential a fourth L2P row.  At
    # N = 8192 (a multiple of 64) the totals stay sorted and are unsorted by
    # their own coalesced passes.
    rng = np.random.default_rng(900)
    K = 32
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    announce("batched potentials", mode, n)
    f = make(hip, x, x, 6, 3, 32, ids=np.arange(n), mode=getattr(hip.HipM2LMode, mode))
    out = f.evaluate_batch_outputs(ms, output="both")
    hb = f.evaluate_batch(ms)
    ref = O.CpuFmm(x, x, 6, 3, 64, fixed_ids=np.arange(n))
    for k in (0, K - 1):
        r = ref.evaluate(ms[k], np.arange(n), output="both")
        e_phi, e_h = rel_l2(out["phi"][k], r["phi"]), rel_l2(out["H"][k], r["H"])
        print(f"  state {k}: rel L2 vs FP64 FMM phi {e_phi:.3e} H {e_h:.3e}", flush=True)
        assert e_phi <= 1e-5 and e_h <= 1e-5
        # Fields with and without the potential rows agree to FP32 rounding.
        assert rel_l2(out["H"][k], hb[k]) <= 1e-6
    only = f.evaluate_batch_outputs(ms, output="potential")
    assert only["H"] is None
    assert rel_l2(only["phi"], out["phi"]) <= 1e-6


@pytest.mark.parametrize("n", [6000, 8192])
def test_device_resident_potentials(hip, n):
    import torch

    rng = np.random.default_rng(910)
    K = 32
    x = rng.uniform(0, 1, size=(n, 3))
    ms = rng.normal(size=(K, n, 3))
    announce("device potentials", n)
    f = make(hip, x, x, 4, 3, 32, ids=np.arange(n), mode=hip.HipM2LMode.MFMA_BF16X3)
    m_dev = torch.tensor(ms, dtype=torch.float32, device="cuda")
    phi = torch.empty((K, n), dtype=torch.float32, device="cuda")
    H = torch.empty((K, n, 3), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    f.evaluate_potential_device(m_dev.data_ptr(), H.data_ptr(), phi.data_ptr(), K, stream)
    phi_only = torch.empty((K, n), dtype=torch.float32, device="cuda")
    f.evaluate_potential_device(m_dev.data_ptr(), 0, phi_only.data_ptr(), K, stream)
    torch.cuda.synchronize()
    host = f.evaluate_batch_outputs(ms, output="both")
    assert rel_l2(phi.cpu().numpy(), host["phi"]) <= 1e-6
    assert rel_l2(H.cpu().numpy(), host["H"]) <= 1e-6
    assert rel_l2(phi_only.cpu().numpy(), host["phi"]) <= 1e-6
