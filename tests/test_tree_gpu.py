"""Device-side leaf order of the uniform tree (needs one MI355X): the stable
sort by leaf Morton index and the leaf ranges, against the host tree build
(the restatement of src/uniform_tree.cpp that every plan uses), bit for bit.
"""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import cdfmm_hip

    assert cdfmm_hip.hip_full_available(), "no gfx950 device is visible to the HIP runtime"
    return cdfmm_hip


def host_tree(hip, x, depth, centre=None, half_width=None):
    o = hip.UniformTreeOptions()
    o.max_level = depth
    if centre is not None:
        o.root_centre = hip.Vec3(*centre)
        o.root_half_width = half_width
    return hip.UniformTree(x, x, o)


def check_against_host(hip, x, depth, centre=None, half_width=None):
    t = host_tree(hip, x, depth, centre, half_width)
    c = t.root_centre
    perm, ranges = hip.device_leaf_order(x, (c.x, c.y, c.z), t.root_half_width, depth)
    np.testing.assert_array_equal(perm, t.source_permutation)
    leaf0 = ((1 << (3 * depth)) - 1) // 7
    np.testing.assert_array_equal(ranges, t.node_ranges[leaf0:, 0:2])
    return t


@pytest.mark.parametrize("n,depth", [(200000, 5), (5000, 3), (777, 0), (1, 2)])
def test_leaf_order_equals_host_tree(hip, n, depth):
    rng = np.random.default_rng(610 + depth)
    x = rng.normal(size=(n, 3)) * [1.0, 2.0, 0.5] + [3.0, -1.0, 0.25]
    print(f"case: leaf order n={n} depth={depth}", flush=True)
    check_against_host(hip, x, depth)


def test_leaf_order_on_cell_faces_and_repeated_points(hip):
    # Points on the grid planes and on the root box's faces (clamped into
    # the last cell), with many points repeated: ties keep the input order.
    rng = np.random.default_rng(620)
    depth = 4
    grid = rng.integers(0, (1 << depth) + 1, size=(3000, 3)) / (1 << depth)
    x = np.concatenate([grid, grid[:500], rng.uniform(0, 1, size=(2000, 3))])
    print("case: leaf order on faces", flush=True)
    check_against_host(hip, x, depth, centre=(0.5, 0.5, 0.5), half_width=0.5)


def test_leaf_order_refuses_points_outside_the_root(hip):
    x = np.array([[0.5, 0.5, 0.5], [1.2, 0.5, 0.5]])
    with pytest.raises(ValueError, match="outside"):
        hip.device_leaf_order(x, (0.5, 0.5, 0.5), 0.5, 3)
    perm, ranges = hip.device_leaf_order(np.zeros((0, 3)), (0.5, 0.5, 0.5), 0.5, 2)
    assert perm.shape == (0,) and ranges.shape == (64, 2) and np.all(ranges == 0)


def test_leaf_order_time_at_one_million_points(hip):
    rng = np.random.default_rng(314160)
    x = rng.uniform(0, 1, size=(1 << 20, 3))
    hip.device_leaf_order(x[:1000], (0.5, 0.5, 0.5), 0.5, 5)  # first-use setup
    t0 = time.perf_counter()
    t = host_tree(hip, x, 5)
    t1 = time.perf_counter()
    c = t.root_centre
    perm, _ = hip.device_leaf_order(x, (c.x, c.y, c.z), t.root_half_width, 5)
    t2 = time.perf_counter()
    print(f"  1M points, depth 5: host tree build {1e3 * (t1 - t0):.1f} ms, device leaf order "
          f"(with transfers) {1e3 * (t2 - t1):.1f} ms", flush=True)
    np.testing.assert_array_equal(perm, t.source_permutation)


def test_leaf_order_through_the_c_abi_on_device_memory(hip):
    # hipfmm_tree_leaf_order with device operands (torch tensors), ordered on
    # the caller's stream.
    import ctypes
    import os

    import torch

    from conftest import REPO

    lib = ctypes.CDLL(os.path.join(REPO, "dip-fmm_amd", "libhipfmm.so"))
    lib.hipfmm_tree_leaf_order.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
    lib.hipfmm_last_error.restype = ctypes.c_char_p
    rng = np.random.default_rng(630)
    n, depth = 50000, 4
    x = rng.uniform(-2.0, 3.0, size=(n, 3))
    t = host_tree(hip, x, depth)
    c = t.root_centre
    centre = (ctypes.c_double * 3)(c.x, c.y, c.z)
    pos = torch.tensor(x, dtype=torch.float64, device="cuda")
    perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    ranges = torch.full((1 << (3 * depth), 2), -1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    print("case: leaf order through the C-ABI, device memory", flush=True)
    rc = lib.hipfmm_tree_leaf_order(pos.data_ptr(), n, centre, t.root_half_width, depth, 1, perm.data_ptr(),
                                    ranges.data_ptr(), ctypes.c_void_p(stream))
    assert rc == 0, lib.hipfmm_last_error().decode()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(perm.cpu().numpy(), t.source_permutation)
    leaf0 = ((1 << (3 * depth)) - 1) // 7
    np.testing.assert_array_equal(ranges.cpu().numpy(), t.node_ranges[leaf0:, 0:2])
    rc = lib.hipfmm_tree_leaf_order(pos.data_ptr(), n, centre, t.root_half_width, 10, 1, perm.data_ptr(),
                                    ranges.data_ptr(), ctypes.c_void_p(stream))
    assert rc == 2 and b"max_level" in lib.hipfmm_last_error()
