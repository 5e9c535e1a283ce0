"""Write the golden fixtures of the reference CpuStatic path.

Run in the build container, where the reference compiles from its own sources
(oracle/Makefile target ``ref``):

    make -C oracle ref && python tests/golden/make_golden.py

Every fixture holds inputs and the compiled reference's outputs:

* ``ref_fields.npz``: CpuStatic fields (FP64 and FP32 plans) for seeded
  uniform clouds -- N=20,000, p=4, depth 4 (the reference's CPU config) and
  N=4,096 at p in {4, 6, 8}, depths 3 and 4 -- targets = sources with the
  identity map, plus one separate-target case without identities.
* ``ref_trees.npz``: permutations, node ranges and list1/list2 (CSR) for
  small edge-case geometries (points on root faces and internal boundaries,
  empty leaves, separate targets, depths 0-3, a user-fixed root).
* ``ref_m2l.npz``: SHA-256 digests of every M2L class matrix at p = 4, 6, 8
  and a few full matrices.
* ``ref_potentials.npz``: CpuStatic potentials and fields (output 'both',
  FP64 and FP32 plans) for the N=20,000 / p=4 / depth 4 cloud, an N=4,096,
  p=8, depth 3 cloud and the separate-target case, plus a cloud with a
  coincident non-identity pair.

Inputs are stored with the outputs, so the fixtures do not depend on NumPy's
generator streams.
"""

from __future__ import annotations

import hashlib
import itertools
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402


def uniform_cloud(n: int, seed: int):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, size=(n, 3))
    m = rng.normal(size=(n, 3))
    return x, m


def fields():
    out = {}
    cases = [("c1", 20000, 4, 4, 314160)]
    for p in (4, 6, 8):
        for d in (3, 4):
            cases.append((f"n4096_p{p}_d{d}", 4096, p, d, 1000 + 10 * p + d))
    for name, n, p, d, seed in cases:
        x, m = uniform_cloud(n, seed)
        if name == "c1":
            m = m / np.linalg.norm(m, axis=1, keepdims=True)
        ids = np.arange(n, dtype=np.int32)
        out[f"{name}/positions"] = x
        out[f"{name}/moments"] = m
        out[f"{name}/config"] = np.array([n, p, d], dtype=np.int64)
        for prec in (64, 32) if name == "c1" or p == 4 else (32,):
            t = time.time()
            f = O.CpuFmm(x, x, p, d, prec, fixed_ids=ids, kind="ref")
            r = f.evaluate(m, ids, output="field")
            out[f"{name}/H{prec}"] = r["H"].astype(np.float64 if prec == 64 else np.float32)
            print(f"{name} fp{prec}: {time.time() - t:.1f}s", flush=True)
    # Separate targets, no identity map, explicit root.
    rng = np.random.default_rng(77)
    s = rng.uniform(-0.9, 0.9, size=(3000, 3))
    t = rng.uniform(-0.9, 0.9, size=(2000, 3))
    m = rng.normal(size=(3000, 3))
    for prec in (64, 32):
        f = O.CpuFmm(s, t, 4, 3, prec, root_centre=(0.0, 0.0, 0.0), root_half_width=1.0, kind="ref")
        out[f"separate/H{prec}"] = f.evaluate(m, output="field")["H"]
    out["separate/sources"] = s
    out["separate/targets"] = t
    out["separate/moments"] = m
    np.savez_compressed(os.path.join(HERE, "ref_fields.npz"), **out)


TREE_ARRAYS = ("source_permutation", "target_permutation", "node_ranges", "list1", "list2")


def tree_cases():
    rng = np.random.default_rng(2024)
    faces = np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [0.5, -0.5, 0.25],
                      [1.0, 0.0, -1.0], [-0.25, 0.75, 0.5], [0.25, 0.25, 0.25]])
    return {
        "faces_d3": (faces, faces, 3, ((0.0, 0.0, 0.0), 1.0)),
        "clustered_d3": (rng.normal(0.3, 0.02, size=(200, 3)), rng.normal(-0.2, 0.05, size=(150, 3)), 3, None),
        "single_d0": (np.array([[0.1, 0.2, 0.3]]), np.array([[0.1, 0.2, 0.3]]), 0, None),
        "uniform_d2": (rng.uniform(0, 1, size=(500, 3)), rng.uniform(0, 1, size=(300, 3)), 2, None),
        "slab_d3": (rng.uniform(0, 1, size=(400, 3)) * [1, 1, 0.1], rng.uniform(0, 1, size=(100, 3)) * [1, 1, 0.1] + [0, 0, 0.2], 3, None),
    }


def trees():
    out = {}
    for name, (s, t, d, root) in tree_cases().items():
        kw = {} if root is None else {"root_centre": root[0], "root_half_width": root[1]}
        f = O.CpuFmm(s, t, 2, d, 64, kind="ref", **kw)
        out[f"{name}/sources"] = s
        out[f"{name}/targets"] = t
        out[f"{name}/depth"] = np.array(d)
        out[f"{name}/root"] = f.root()
        if root is not None:
            out[f"{name}/fixed_root"] = np.array([*root[0], root[1]])
        for a in TREE_ARRAYS:
            out[f"{name}/{a}"] = f.array(a)
    np.savez_compressed(os.path.join(HERE, "ref_trees.npz"), **out)


def m2l():
    out = {}
    offsets = [o for o in itertools.product(range(-3, 4), repeat=3) if max(abs(v) for v in o) >= 2]
    for p in (4, 6, 8):
        t = time.time()
        digests = []
        for o in offsets:
            mat = O.m2l_matrix(p, o, kind="ref")
            digests.append(hashlib.sha256(np.ascontiguousarray(mat).tobytes()).hexdigest())
            if o in ((2, 0, 0), (3, -2, 1), (-3, -3, -3)):
                out[f"p{p}/matrix_{o[0]}_{o[1]}_{o[2]}"] = mat
        out[f"p{p}/offsets"] = np.array(offsets, dtype=np.int64)
        out[f"p{p}/sha256"] = np.array(digests)
        print(f"m2l p{p}: {time.time() - t:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_m2l.npz"), **out)


def potentials():
    out = {}
    cases = [("c1", 20000, 4, 4, 314160), ("n4096_p8_d3", 4096, 8, 3, 1083)]
    for name, n, p, d, seed in cases:
        x, m = uniform_cloud(n, seed)
        if name == "c1":
            m = m / np.linalg.norm(m, axis=1, keepdims=True)
        ids = np.arange(n, dtype=np.int32)
        out[f"{name}/positions"] = x
        out[f"{name}/moments"] = m
        out[f"{name}/config"] = np.array([n, p, d], dtype=np.int64)
        for prec in (64, 32):
            t = time.time()
            f = O.CpuFmm(x, x, p, d, prec, fixed_ids=ids, kind="ref")
            r = f.evaluate(m, ids, output="both")
            out[f"{name}/phi{prec}"] = r["phi"]
            out[f"{name}/H{prec}"] = r["H"]
            print(f"potential {name} fp{prec}: {time.time() - t:.1f}s", flush=True)
    rng = np.random.default_rng(77)
    s = rng.uniform(-0.9, 0.9, size=(3000, 3))
    t = rng.uniform(-0.9, 0.9, size=(2000, 3))
    m = rng.normal(size=(3000, 3))
    for prec in (64, 32):
        f = O.CpuFmm(s, t, 4, 3, prec, root_centre=(0.0, 0.0, 0.0), root_half_width=1.0, kind="ref")
        r = f.evaluate(m, output="both")
        out[f"separate/phi{prec}"] = r["phi"]
        out[f"separate/H{prec}"] = r["H"]
    out["separate/sources"] = s
    out["separate/targets"] = t
    out["separate/moments"] = m
    # Coincident non-identity pair: points 0 and 1 coincide, target 1 has no
    # identity, so its potential (and field) is NaN.
    x = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.8, 0.7, 0.6], [0.4, 0.5, 0.9]])
    m = np.eye(4, 3) + 0.1
    for prec in (64, 32):
        f = O.CpuFmm(x, x, 2, 1, prec, kind="ref")
        r = f.evaluate(m, np.array([0, -1, 2, 3], dtype=np.int32), output="both")
        out[f"coincident/phi{prec}"] = r["phi"]
    out["coincident/positions"] = x
    out["coincident/moments"] = m
    np.savez_compressed(os.path.join(HERE, "ref_potentials.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref")
    which = sys.argv[1:] or ["trees", "fields", "m2l", "potentials"]
    for w in which:
        globals()[w]()
