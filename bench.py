#!/usr/bin/env python3
"""HipFull headline benchmark (one MI355X).

Metric: field evaluations per second at 2^20 dipoles, p = 8, K = 32 moment
states -- the MFMA-batched evaluate, reported with the plain-HIP FP32 rate
beside it and the accuracy of the MFMA result.

A step is one batched evaluate of K = 32 moment states on the fixed plan
(2^20 dipoles uniform in [0,1)^3, targets = sources with the identity map,
expansion order 8, uniform depth 5, FP32 plan, M2L on the matrix cores with
bf16 hi/lo split operands).  Moments and fields stay in HBM (torch tensors);
value = N * K * steps / elapsed seconds.  The plain path (M2L on the vector
ALUs in FP32, one moment state per evaluate) is timed on the same plan.

Accuracy: rel. L2 of the MFMA result against the FP64 FMM at the same (p,
depth) (the HipFull FP64 plan, itself checked against CpuStatic to 1e-12 by
the tests) and against the FP64 direct sum on 4096 seeded sample targets.

Roofline: the dominant kernel is the M2L product (one launch over levels
2..5); its algorithmic work is 2 nc^2 flops per list-2 interaction and moment
state, nc = 81 at p = 8 (the class matrices factor through the (p+1)^2-
dimensional harmonic subspace; the reference's form uses n = 165 and is
reported beside it), timed with HIP events on the plan's stream around that
launch.

CPU baseline: the repository's CpuStatic restatement (oracle/, "port") on a
bounded sample of the same workload -- 2^17 dipoles at the same density
(32 per leaf, depth 4), p = 8, one FP32 moment state -- timed on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "dip-fmm_amd"))

MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed)
SEED = 314160                   # the reference benchmark seed 314159 + config index 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def uniform_geometry(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0.0, 1.0, size=(n, 3))


def moment_states(n: int, k: int, seed: int) -> np.ndarray:
    out = np.empty((k, n, 3), dtype=np.float32)
    for i in range(k):
        out[i] = np.random.default_rng(seed + 1000 + i).standard_normal((n, 3), dtype=np.float32)
    return out


def make_plan(hip, x, order, depth, precision, mode):
    o = hip.UniformFmmOptions()
    o.expansion_order = order
    o.tree.max_level = depth
    o.backend = hip.ExecutionBackend.HIP_FULL
    o.precision = precision
    o.fixed_target_source_indices = list(range(len(x)))
    o.hip_m2l_mode = mode
    return hip.UniformFmm(x, x, o)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def measured_traffic(n: int, order: int, depth: int, k: int):
    """HBM bytes per M2L launch at the headline workload, from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB copy);
    None for any other workload."""
    path = os.path.join(HERE, "profiles", "m2l_mfma_1M_p8_d5_K32_traffic.json")
    if (n, order, depth, k) != (1 << 20, 8, 5, 32) or not os.path.exists(path):
        return None
    with open(path) as fh:
        return float(json.load(fh)["traffic_bytes_per_launch"])


def cpu_baseline(order: int) -> dict:
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle

    n, depth = 1 << 17, 4
    x = uniform_geometry(n, SEED)
    m = np.random.default_rng(SEED + 1000).standard_normal((n, 3))
    ids = np.arange(n)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    f = oracle.CpuFmm(x, x, order, depth, 32, fixed_ids=ids)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    f.evaluate(m)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: setup {setup:.1f}s, evaluate {dt:.2f}s on {threads} threads")
    return {
        "value": n / dt,
        "unit": "field evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CpuStatic restatement, FP32, 2^17 uniform dipoles (32 per leaf, depth {depth}), p={order}, "
                  f"1 moment state, one evaluate after plan setup ({setup:.0f}s setup, {dt:.2f}s evaluate)",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--order", type=int, default=8)
    ap.add_argument("--depth", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-split M2L line")
    ap.add_argument("--autotune", action="store_true",
                    help="run the plan's launch autotuner at K states before timing")
    ap.add_argument("--near-variant", type=int, default=0,
                    help="pin the near-field kernel (codes of hip_plan_statistics; 0 = built-in default)")
    ap.add_argument("--no-other-paths", action="store_true",
                    help="skip the adjoint, potential and field-gradient timings")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB before the steps (the reference kernel "
                         "for scaling PMC byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the HipFull evaluate runs on a single MI355X and this build has no "
              "multi-GPU path; run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    import cdfmm_hip as hip

    if not hip.hip_full_available():
        raise RuntimeError("HipFull needs a gfx950 device")
    dev = torch.device("cuda:0")
    N, K, p, depth = args.n, args.batch, args.order, args.depth
    log(f"device: {hip.hip_device_description()}")
    t0 = time.perf_counter()
    x = uniform_geometry(N, SEED)
    fmm = make_plan(hip, x, p, depth, hip.StaticPrecision.FLOAT32, hip.HipM2LMode.MFMA_BF16X3)
    if args.near_variant:
        fmm.set_launch_choices(0, args.near_variant)
    if args.autotune:
        fmm.autotune(K)
        st = fmm.hip_plan_statistics
        log(f"autotune: M2L tile shape {st['m2l_tile_shape']}, near-field variant {st['p2p_kernel_variant']} "
            f"in {st['autotune_seconds']:.1f}s")
    log(f"plan setup: {time.perf_counter() - t0:.1f}s")
    stats = fmm.hip_plan_statistics
    moments = torch.from_numpy(moment_states(N, K, SEED)).to(dev)
    fields = torch.empty((K, N, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        fmm.evaluate_device(moments.data_ptr(), fields.data_ptr(), K, stream)

    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written")
        del src, dst

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # Every M2L launch is timed by its own pair of HIP events on the plan's
    # stream; the difference of the running totals around the timed region
    # gives that region's average launch time without synchronising steps.
    before = fmm.device_timings()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    value = N * K * args.steps / elapsed
    stage = fmm.device_timings()
    launches = stage["m2l_kernel_launches"] - before["m2l_kernel_launches"]
    if launches != args.steps:
        raise RuntimeError(f"expected {args.steps} M2L launches in the timed region, saw {launches}")
    m2l_kernel = [(stage["m2l_kernel_total"] - before["m2l_kernel_total"]) / launches]
    log(f"MFMA K={K}: {ms_per_step:.2f} ms/step, {value:.3e} evals/s; stages {stage}")

    pairs = stats["m2l_interaction_count"]
    n = fmm.coefficient_count
    nc = int(stats["m2l_operand_size"])  # (p+1)^2 when the class matrices factor through the harmonic subspace
    m2l_flops = 2.0 * nc * nc * pairs * K
    m2l_s = float(np.mean(m2l_kernel))
    achieved = m2l_flops / m2l_s / 1e12
    roofline = {
        "bound": "mfma",
        "achieved": achieved,
        "peak": MFMA_BF16_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / MFMA_BF16_PEAK_TFLOPS,
        "traffic": measured_traffic(N, p, depth, K),
        "kernel": "m2l_mfma_kernel (one launch, levels 2..L)",
        "kernel_ms": 1e3 * m2l_s,
        "algorithmic_flops_per_launch": m2l_flops,
        "m2l_operand_size": nc,
        "reference_equivalent_flops_per_launch": 2.0 * n * n * pairs * K,
        "list2_interactions": pairs,
        "split_products_per_flop": 3,
    }

    result_fields = fields[0].float().cpu().numpy()
    plain = {}
    if not args.no_plain:
        fmm.set_m2l_mode(hip.HipM2LMode.VALU)
        one_m = moments[0:1].contiguous()
        one_f = torch.empty((1, N, 3), dtype=torch.float32, device=dev)
        fmm.evaluate_device(one_m.data_ptr(), one_f.data_ptr(), 1, stream)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain_steps):
            fmm.evaluate_device(one_m.data_ptr(), one_f.data_ptr(), 1, stream)
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - tp) / args.plain_steps
        pt = fmm.device_timings()
        plain = {"value": N / plain_s, "ms_per_evaluate": 1e3 * plain_s, "m2l_kernel_ms": 1e3 * pt["m2l_kernel"]}
        plain_fields = one_f[0].cpu().numpy()
        fmm.set_m2l_mode(hip.HipM2LMode.MFMA_BF16X3)
        log(f"plain VALU K=1: {1e3 * plain_s:.2f} ms/evaluate, {N / plain_s:.3e} evals/s")

    # The same step with the level-scaled FP16 hi/lo M2L (same MFMA rate,
    # 22 bits per operand instead of 16), reported beside the headline.
    fp16 = {}
    fp16_fields = None
    if not args.no_fp16:
        fmm.set_m2l_mode(hip.HipM2LMode.MFMA_FP16X3)
        step()
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 5))
        t16 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        s16 = (time.perf_counter() - t16) / reps
        fp16 = {"value": N * K / s16, "ms_per_step": 1e3 * s16}
        fp16_fields = fields[0].float().cpu().numpy()
        fmm.set_m2l_mode(hip.HipM2LMode.MFMA_BF16X3)
        log(f"MFMA FP16x3 K={K}: {1e3 * s16:.2f} ms/step, {N * K / s16:.3e} evals/s")

    # The other device paths of the same plan at the same workload: the
    # adjoint (cotangents in, moment gradients out), fields with potentials,
    # and the field gradient (6 entries per target).
    other = {}
    if not args.no_other_paths:
        reps = max(1, min(args.steps, 3))

        def time_path(run):
            run()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / reps

        grad = torch.empty_like(moments)
        sec = time_path(lambda: fmm.evaluate_adjoint_device(fields.data_ptr(), grad.data_ptr(), K, stream))
        other["adjoint"] = {"ms_per_step": 1e3 * sec, "value": N * K / sec}
        del grad
        pot = torch.empty((K, N), dtype=torch.float32, device=dev)
        sec = time_path(lambda: fmm.evaluate_potential_device(moments.data_ptr(), fields.data_ptr(), pot.data_ptr(),
                                                              K, stream))
        other["fields_and_potentials"] = {"ms_per_step": 1e3 * sec, "value": N * K / sec}
        del pot
        gmat = torch.empty((K, N, 6), dtype=torch.float32, device=dev)
        sec = time_path(lambda: fmm.evaluate_field_gradient_device(moments.data_ptr(), gmat.data_ptr(), K, stream))
        other["field_gradient"] = {"ms_per_step": 1e3 * sec, "value": N * K / sec}
        del gmat
        torch.cuda.empty_cache()
        log("other paths (ms per K-state step): " + ", ".join(f"{k} {v['ms_per_step']:.2f}" for k, v in other.items()))

    accuracy = {}
    if not args.no_accuracy:
        del moments, fields
        torch.cuda.empty_cache()
        f64 = make_plan(hip, x, p, depth, hip.StaticPrecision.FLOAT64, hip.HipM2LMode.VALU)
        m0 = moment_states(N, 1, SEED)[0].astype(np.float64)
        ref64 = f64.evaluate(m0)["H"]
        accuracy["mfma_vs_fp64_fmm"] = rel_l2(result_fields, ref64)
        if plain:
            accuracy["plain_fp32_vs_fp64_fmm"] = rel_l2(plain_fields, ref64)
        if fp16_fields is not None:
            accuracy["mfma_fp16x3_vs_fp64_fmm"] = rel_l2(fp16_fields, ref64)
        del f64
        sample = np.random.default_rng(SEED + 7).choice(N, size=4096, replace=False)
        import ctypes

        lib = ctypes.CDLL(os.path.join(HERE, "dip-fmm_amd", "libhipfmm.so"))
        dp = ctypes.POINTER(ctypes.c_double)
        tgt = np.ascontiguousarray(x[sample])
        ids = np.ascontiguousarray(sample.astype(np.int32))
        direct = np.zeros((4096, 3))
        rc = lib.hipfmm_direct_field(tgt.ctypes.data_as(dp), 4096, x.ctypes.data_as(dp), N, m0.ctypes.data_as(dp),
                                     ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), direct.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError("GPU direct sum failed")
        accuracy["mfma_vs_fp64_direct_4096"] = rel_l2(result_fields[sample], direct)
        accuracy["fp64_fmm_vs_fp64_direct_4096"] = rel_l2(ref64[sample], direct)
        log(f"accuracy: {accuracy}")

    baseline = None if args.no_cpu_baseline else cpu_baseline(p)

    line = {
        "metric": "field evals/s, 1M dipoles p=8 K=32: MFMA-batched vs plain HIP @ rel err≤1e-5",
        "value": value,
        "unit": "field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: uniform [0,1)^3 positions, N(0,1) moment components, seeded",
        "config": {
            "workload": f"{N} dipoles, targets = sources (identity map), p={p}, uniform depth {depth}, "
                        f"K={K} moment states per evaluate, FP32 plan, M2L on MFMA (bf16 hi/lo x3, fp32 accumulate)",
            "n": N, "order": p, "depth": depth, "batch": K, "m2l": "mfma_bf16x3",
        },
        "plain_hip_fp32": plain or None,
        "mfma_fp16x3": fp16 or None,
        "other_paths": other or None,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "accuracy": accuracy or None,
        "stages_ms": {k: 1e3 * v for k, v in stage.items() if k not in ("evaluations", "m2l_kernel_launches")},
        "roofline": roofline,
        "cpu_baseline": baseline,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
