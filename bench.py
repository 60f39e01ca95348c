#!/usr/bin/env python3
"""Throughput of the MI355X ICON core on the batched 512^2 tilt-series workload.

One step = one batched ICON solve of every slice of the synthetic tilt series
(256 slices of 512 x 512, 41 tilts +-60 deg in 3 deg steps, 200 iterations
split 20 INFR / 160 ICON / 20 INFR, threshold m >= 0) -- the whole solve
replayed as one HIP graph with the sinograms already resident in HBM.  The
data residual R(m) = sqrt(sum w|Am-b|^2 / sum w|b|^2) (FP64) of every slice is
produced inside the timed region.

Prints one JSON line with the throughput (slices/s), the roofline of the
dominant kernel (algorithmic bytes / HIP-event-timed launch duration against
8 TB/s HBM) and the CPU baseline (the FP32-mirror oracle on the host cores on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "512² slices/s to fixed data residual, learned-start vs plain ICON, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(plan):
    """Per-slice algorithmic bytes of each kernel (c64 = 8 B, f32 = 4 B); node
    coordinates, window tables and weights are shared by all slices and left out."""
    M, n0, n1, N0, N1 = plan.M, plan.n0, plan.n1, plan.N0, plan.N1
    grid = 8 * n0 * n1
    return {
        # interpolation: the full oversampled grid once + the node values once
        "col_fwd": grid + 8 * M,
        # spread: the node values once + the full oversampled grid once
        "col_adj": 8 * M + grid,
        # image in (complex gradient) + the N0 non-zero grid rows out
        "row_fwd": 8 * N0 * N1 + 8 * N0 * n1,
        # N0 grid rows in + A^H W b in + gradient out
        "row_adj": 8 * N0 * n1 + 16 * N0 * N1,
        # interpolation partials in (<= 2 per node) + data + weighted values out
        "combine": 16 * M + 8 * M + 8 * M,
    }


LAUNCHES_PER_ITER = {"row_fwd": 2, "col_fwd": 2, "combine": 2, "col_adj": 1, "row_adj": 1}


def cpu_baseline(cfg, S_cpu, iters_cpu, total_iters, threads):
    import oracle
    plan = oracle.Plan(cfg["N"], cfg["N"], cfg["tilt"], 32)
    sinos = cfg["sino"][:S_cpu]
    pre = min(cfg["pre"], iters_cpu)
    t0 = time.perf_counter()
    _, used = plan.solve_batch(sinos, pre, iters_cpu - pre, 0, 2, 0.0, threads=threads)
    dt = time.perf_counter() - t0
    value = S_cpu * (iters_cpu / total_iters) / dt
    return dict(value=round(value, 4), unit="slices/s", cores=int(used), kind="port",
                sample="%d slices x %d of %d iterations of the same 512^2 x %d-tilt series, FP32-mirror "
                       "oracle (OpenMP over slices); slices/s scaled by %d/%d; %.1f s wall"
                       % (S_cpu, iters_cpu, total_iters, len(cfg["tilt"]), iters_cpu, total_iters, dt))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--iters", default="20,160,20", help="pre,icon,post")
    ap.add_argument("--cpu-slices", type=int, default=16)
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--col-block", type=int, default=0, help="grid columns per spread/interp tile (0: default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the ICON solve shards by slice with no exchange between devices, "
              "and this build runs on a single MI355X; --gpus %d is not supported." % args.gpus)
        sys.exit(2)

    import feilab_amd
    from feilab_amd import phantom
    from feilab_amd.icon import DeviceArray, SolveOptions

    pre, icon_it, post = [int(v) for v in args.iters.split(",")]
    iters = pre + icon_it + post
    S = args.slices
    log("generating the synthetic tilt series: %d slices of 512^2, 41 tilts" % S)
    t0 = time.perf_counter()
    cfg = phantom.config2(nslices=S)
    log("data ready in %.1f s" % (time.perf_counter() - t0))

    plan = feilab_amd.Plan(cfg["N"], cfg["tilt"], max_slices=S, device=0)
    if args.col_block:
        plan.set_col_block(args.col_block)
    d_sino = DeviceArray.from_host(cfg["sino"], np.float32)
    d_m = DeviceArray((S, plan.N0, plan.N1), np.float32)
    d_res = DeviceArray((S, iters + 1), np.float64)
    opts = SolveOptions(pre, icon_it, post, 2, 0.0, 1)
    L = feilab_amd.lib()

    def step():
        plan.reconstruct_dev(S, d_sino, d_m, opts, d_resid=d_res)

    for w in range(args.warmup):
        step()
        L.icon_device_synchronize()
        log("warmup step %d done" % w)
    L.icon_device_synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
    L.icon_device_synchronize()
    dt = time.perf_counter() - t0
    log("timed %d steps in %.3f s" % (args.steps, dt))
    ms_step = dt / args.steps * 1e3
    value = S * args.steps / dt
    resid = d_res.numpy()
    final_res = resid[:, -1]

    # per-kernel launch durations (HIP events on the plan's stream)
    kb = algorithmic_bytes(plan)
    kt = {k: plan.time_kernel(k, S, args.kernel_reps) for k in LAUNCHES_PER_ITER}
    share = {k: kt[k] * LAUNCHES_PER_ITER[k] for k in kt}
    per_iter_ms = sum(share.values())
    kernels = {}
    for k in kt:
        gbs = kb[k] * S / (kt[k] * 1e-3) / 1e9
        kernels[k] = dict(ms=round(kt[k], 4), algo_bytes_per_slice=kb[k], GBs=round(gbs, 1),
                          share=round(share[k] / per_iter_ms, 3))
    log("kernels:", json.dumps(kernels))
    dom = max(("col_fwd", "col_adj"), key=lambda k: share[k])
    ach = kb[dom] * S / (kt[dom] * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                    algo_bytes="%d per slice x %d slices per launch" % (kb[dom], S))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline: %d slices x %d iterations" % (args.cpu_slices, args.cpu_iters))
        cpu = cpu_baseline(cfg, args.cpu_slices, args.cpu_iters, iters, min(args.cpu_threads, os.cpu_count() or 1))
        log("cpu baseline done")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched plain ICON: %d slices of 512x512, 41 tilts +-60 deg step 3, %d iterations "
                        "(%d,%d,%d), m >= 0 threshold, all slices per launch, whole solve as one HIP graph"
                        % (S, iters, pre, icon_it, post),
            "slices": S, "N": 512, "tilts": 41, "iterations": iters, "start": "plain (zeros)",
            "fixed_residual": "R after %d plain iterations" % iters,
            "final_residual_mean": float(final_res.mean()), "final_residual_max": float(final_res.max()),
            "col_block": int(plan.info.col_block),
            "ms_per_iteration_kernels": round(per_iter_ms, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
