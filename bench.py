#!/usr/bin/env python3
"""Throughput of the MI355X ICON core on the batched 512^2 tilt-series workload:
slices/s to a fixed data residual, learned start against plain ICON.

Workload: every slice of a synthetic tilt series (256 slices of 512 x 512,
41 tilts +-60 deg in 3 deg steps, SNR 0.05), sinograms resident in HBM.

* Plain leg: one step = the batched plain ICON solve (200 iterations split
  20 INFR / 160 ICON / 20 INFR, threshold m >= 0) replayed as one HIP graph.
  Its per-slice data residual R(m) = sqrt(sum w|Am-b|^2 / sum w|b|^2) (FP64)
  is the target, its reprojection FRC at the min-|theta| tilt the reference.
* Learned leg: a U-Net trained on the device (its own BF16 MFMA kernels) on
  (4-iteration, 200-iteration) pairs of a separate phantom; one step = 4
  plain iterations, the U-Net warm start, and k descent iterations, with k the
  smallest continuation that brings every slice to R <= its plain target with
  the reprojection FRC (rings 1..N/4) within 0.005 of plain ICON's.  Slices
  that miss the target are re-solved by plain ICON (counted as fallbacks).

Prints one JSON line: value = learned-start slices/s, plus the plain leg,
the roofline of the dominant ICON kernel (algorithmic bytes / HIP-event-timed
launch duration against 8 TB/s HBM) and the CPU baseline (the FP32-mirror
oracle on the host cores on a bounded sample of the plain workload).
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

# U-Net convolutions (Cin, Cout, side at a 512^2 input) for the forward FLOP count
UNET_CONVS = [(1, 32, 512), (32, 32, 512), (32, 64, 256), (64, 64, 256), (64, 128, 128), (128, 128, 128),
              (128, 256, 64), (256, 256, 64), (384, 128, 128), (128, 128, 128), (192, 64, 256), (64, 64, 256),
              (96, 32, 512), (32, 32, 512)]
BF16_DENSE_PEAK_TFLOPS = 2500.0


def unet_gflop_per_slice():
    """Forward FLOPs of the U-Net on one 512^2 slice: 3x3 convolutions and the
    1x1 head, 2 FLOP per multiply-add (pool, upsample and ReLU not counted)."""
    return (sum(2.0 * h * h * ci * co * 9 for ci, co, h in UNET_CONVS) + 2.0 * 512 * 512 * 32) / 1e9
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
        # image in (the update reads the gradient and rewrites the image) + the N0 non-zero grid rows out
        "row_fwd": 8 * N0 * N1 + 8 * N0 * n1,
        # N0 grid rows in + A^H W b in + gradient out + the N0 grid rows of A rhat out
        "row_adj": 8 * N0 * n1 + 16 * N0 * N1 + 8 * N0 * n1,
        # interpolation partials in (<= 2 per node) + data + weighted values out
        "combine": 16 * M + 8 * M + 8 * M,
    }


# row_adj also does the row pass of A rhat (fused), so row_fwd runs once
LAUNCHES_PER_ITER = {"row_fwd": 1, "col_fwd": 2, "combine": 2, "col_adj": 1, "row_adj": 1}


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


COPY_BYTES = 1 << 30


def iteration_roofline(plan, S, per_iter_ms):
    """One ICON iteration against HBM with the reference's dataflow bytes per
    slice, 17 G + 6 S + 8 I_c + 3 I_r (G = 8 n0 n1 grid, S = 8 M samples,
    I_c = 8 N0 N1 complex image, I_r = 4 N0 N1 real image: three FFTs, pads,
    spread, two interpolations, weights, residual and update as separate
    passes).  Fusion moves fewer bytes, so the fraction can exceed 1."""
    G = 8 * plan.n0 * plan.n1
    Sb = 8 * plan.M
    Ic = 8 * plan.N0 * plan.N1
    Ir = 4 * plan.N0 * plan.N1
    per_slice = 17 * G + 6 * Sb + 8 * Ic + 3 * Ir
    ach = per_slice * S / (per_iter_ms * 1e-3) / 1e9
    return dict(bytes_per_slice=per_slice, ms_per_iteration=round(per_iter_ms, 4), achieved=round(ach, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4))


def probe(what, S):
    """Child-process mode for the counter passes: run the named ICON kernel on
    the workload's buffers (or a 1 GiB plain copy, or the U-Net forward) a few
    times and exit."""
    import feilab_amd
    from feilab_amd import phantom
    from feilab_amd.icon import DeviceArray, SolveOptions, copy_probe
    if what == "copy":
        copy_probe(COPY_BYTES, 2)
        return
    if what == "unet":
        # the U-Net forward of the learned leg (random weights), twice
        from feilab_amd.unet import UNet
        net = UNet(512, 512, max_batch=min(128, S), seed=1)
        d_in = DeviceArray.from_host(np.random.default_rng(0).standard_normal((S, 512, 512)).astype(np.float32))
        d_out = DeviceArray((S, 512, 512), np.float32)
        for _ in range(2):
            net.warm_start_dev(S, d_in, d_out)
        feilab_amd.lib().icon_device_synchronize()
        net.close()
        return
    cfg = phantom.config2(nslices=S)
    plan = feilab_amd.Plan(cfg["N"], cfg["tilt"], max_slices=S, device=0)
    d_s = DeviceArray.from_host(cfg["sino"], np.float32)
    d_m = DeviceArray((S, plan.N0, plan.N1), np.float32)
    plan.reconstruct_dev(S, d_s, d_m, SolveOptions(1, 1, 0, 2, 0.0, 0, 0))
    feilab_amd.lib().icon_device_synchronize()
    plan.time_kernel(what, S, 2)


def counter_pass(counter, target, S, kernel_substr):
    """Average per-dispatch value (KiB) of one TCC counter for the dispatches
    whose name contains kernel_substr, from a rocprofv3 --pmc child run of the
    probe mode; None if the profiler is unavailable or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="bench_pmc_")
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "p", "--",
           sys.executable, os.path.abspath(__file__), "--probe", target, "--slices", str(S)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120,
                           env=dict(os.environ, TMPDIR=d))
    except subprocess.TimeoutExpired:
        log("counter pass %s on %s timed out" % (counter, target))
        return None
    vals = {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for row in csv.DictReader(open(os.path.join(root, f))):
                    if kernel_substr in row["Kernel_Name"] and row["Counter_Name"] == counter:
                        vals[row["Dispatch_Id"]] = vals.get(row["Dispatch_Id"], 0.0) + float(row["Counter_Value"])
    shutil.rmtree(d, ignore_errors=True)
    if r.returncode != 0 and not vals:
        log("counter pass %s on %s failed (rc %d)" % (counter, target, r.returncode))
        return None
    return float(np.mean(list(vals.values()))) if vals else None


def measure_traffic(kernel, S):
    """HBM bytes per launch of `kernel` from FETCH_SIZE and WRITE_SIZE, each
    scaled by what the same counter reads for a plain copy of known size."""
    name = {"col_fwd": "k_col_fwd", "col_adj": "k_col_adj"}[kernel]
    f_k = counter_pass("FETCH_SIZE", kernel, S, name)
    w_k = counter_pass("WRITE_SIZE", kernel, S, name)
    f_c = counter_pass("FETCH_SIZE", "copy", S, "k_copy_f4")
    w_c = counter_pass("WRITE_SIZE", "copy", S, "k_copy_f4")
    if None in (f_k, w_k, f_c, w_c) or f_c <= 0 or w_c <= 0:
        return None, None
    fs, ws = COPY_BYTES / (f_c * 1024), COPY_BYTES / (w_c * 1024)
    traffic = f_k * 1024 * fs + w_k * 1024 * ws
    return traffic, dict(fetch_kib=round(f_k, 1), write_kib=round(w_k, 1), copy_fetch_kib=round(f_c, 1),
                         copy_write_kib=round(w_c, 1), fetch_scale=round(fs, 4), write_scale=round(ws, 4))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--iters", default="20,160,20", help="pre,icon,post")
    ap.add_argument("--start", type=int, default=4, help="plain iterations before the U-Net")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--train-slices", type=int, default=256)
    ap.add_argument("--unet-batch", type=int, default=128, help="slices per U-Net forward launch sequence")
    ap.add_argument("--max-cont", type=int, default=40, help="continuation iterations tried")
    ap.add_argument("--frc-tol", type=float, default=0.005)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=16)
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--col-block", type=int, default=0, help="grid columns per spread/interp tile (0: default)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-kb", action="store_true", help="skip the Kaiser-Bessel window variant")
    ap.add_argument("--probe", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.probe:
        probe(args.probe, args.slices)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the ICON solve shards by slice with no exchange between devices, "
              "and this build runs on a single MI355X; --gpus %d is not supported." % args.gpus)
        sys.exit(2)

    import feilab_amd
    from feilab_amd import phantom, warmstart
    from feilab_amd.icon import DeviceArray, SolveOptions

    pre, icon_it, post = [int(v) for v in args.iters.split(",")]
    iters = pre + icon_it + post
    S = args.slices
    log("generating the synthetic tilt series: %d slices of 512^2, 41 tilts" % S)
    t0 = time.perf_counter()
    cfg = phantom.config2(nslices=S)
    log("data ready in %.1f s" % (time.perf_counter() - t0))

    plan = feilab_amd.Plan(cfg["N"], cfg["tilt"], max_slices=S, device=0)
    autotune = None
    if args.col_block:
        plan.set_col_block(args.col_block)
    else:
        c, ms = plan.autotune(S)
        autotune = dict(col_block=c, ms_column_kernels_per_iteration={str(k): round(v, 4) for k, v in ms.items()})
        log("autotuned column tile:", json.dumps(autotune))
    iomit = plan.info.iomit
    band = (1, plan.N0 // 4)
    d_sino = DeviceArray.from_host(cfg["sino"], np.float32)
    d_m = DeviceArray((S, plan.N0, plan.N1), np.float32)
    d_res = DeviceArray((S, iters + 1), np.float64)
    d_rp = DeviceArray((S, plan.N0), np.float32)
    opts = SolveOptions(pre, icon_it, post, 2, 0.0, 1, 0)
    L = feilab_amd.lib()

    def timed(step, what):
        for w in range(args.warmup):
            step()
            L.icon_device_synchronize()
            log("%s warmup step %d done" % (what, w))
        L.icon_device_synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step()
        L.icon_device_synchronize()
        dt = time.perf_counter() - t0
        log("%s: timed %d steps in %.3f s" % (what, args.steps, dt))
        return dt

    # ---- plain leg
    dt_plain = timed(lambda: plan.reconstruct_dev(S, d_sino, d_m, opts, d_resid=d_res, d_reproj=d_rp), "plain")
    plain_value = S * args.steps / dt_plain
    R_plain = d_res.numpy()[:, -1]
    frc_plain_curve, frc_plain = warmstart.reprojection_frc(d_rp.numpy(), cfg["sino"], iomit, band)
    log("plain: %.2f slices/s, R mean %.5f, FRC %.4f" % (plain_value, R_plain.mean(), frc_plain))

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
    if not args.no_traffic:
        log("counter passes (FETCH_SIZE, WRITE_SIZE) on %s and a 1 GiB copy" % dom)
        traffic, detail = measure_traffic(dom, S)
        if traffic is not None:
            roofline["traffic"] = int(traffic)
            roofline["traffic_detail"] = detail
            roofline["algo_bytes_per_launch"] = kb[dom] * S
        log("traffic:", roofline["traffic"])

    # ---- Kaiser-Bessel variant: same workload, 6-tap window
    kb6 = None
    if not args.no_kb:
        log("Kaiser-Bessel (width 6) plan: kernel timing and a plain solve")
        pk = feilab_amd.Plan(cfg["N"], cfg["tilt"], max_slices=S, device=0, window="kb6")
        kk = {}
        for k in ("col_fwd", "col_adj"):
            ms = pk.time_kernel(k, S, args.kernel_reps)
            gbs = kb[k] * S / (ms * 1e-3) / 1e9
            kk[k] = dict(ms=round(ms, 4), GBs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        d_mk = DeviceArray((S, plan.N0, plan.N1), np.float32)
        d_rk = DeviceArray((S, iters + 1), np.float64)
        d_pk = DeviceArray((S, plan.N0), np.float32)
        run_k = lambda: pk.reconstruct_dev(S, d_sino, d_mk, opts, d_resid=d_rk, d_reproj=d_pk)
        dt_k = timed(run_k, "plain kb6")
        R_k = d_rk.numpy()[:, -1]
        _, frc_k = warmstart.reprojection_frc(d_pk.numpy(), cfg["sino"], iomit, band)
        m_g, m_k = d_m.numpy(), d_mk.numpy()
        kb6 = dict(window="Kaiser-Bessel, width 6, beta %.4f" % pk.info.b0, kernels=kk,
                   plain_slices_per_s=round(S * args.steps / dt_k, 3),
                   resid_max_over_gauss=round(float((R_k / R_plain).max()), 5),
                   rel_l2_vs_gauss_max=float(np.max(np.linalg.norm((m_k - m_g).reshape(S, -1), axis=1)
                                                    / np.linalg.norm(m_g.reshape(S, -1), axis=1))),
                   frc=round(frc_k, 5))
        log("kb6:", json.dumps(kb6))
        for a in (d_mk, d_rk, d_pk):
            a.free()
        pk.close()

    # ---- learned leg
    learned = None
    unet_roofline = None
    value, ms_step = plain_value, dt_plain / args.steps * 1e3
    if not args.no_learned:
        log("training the warm start: %d slices of a separate phantom, %d steps" % (args.train_slices, args.train_steps))
        t0 = time.perf_counter()
        tr = phantom.config2(nslices=args.train_slices, seed=2001)
        params, losses = warmstart.train_warm_start(plan, tr["sino"], args.start, (pre, icon_it, post),
                                                    steps=args.train_steps, log=log)
        train_s = time.perf_counter() - t0
        solver = warmstart.LearnedSolver(plan, params, S, start=args.start, batch=args.unet_batch)
        d_ml = DeviceArray((S, plan.N0, plan.N1), np.float32)

        def check(k, use_net):
            d_r = DeviceArray((S, k + 1), np.float64)
            solver.run(S, d_sino, d_ml, k, d_resid=d_r, d_reproj=d_rp, use_net=use_net)
            L.icon_device_synchronize()
            R = d_r.numpy()[:, -1]
            _, f = warmstart.reprojection_frc(d_rp.numpy(), cfg["sino"], iomit, band)
            ok = bool((R <= R_plain).all() and abs(f - frc_plain) <= args.frc_tol)
            return ok, R, f

        k_star, res = None, None
        for k in range(1, args.max_cont + 1):
            ok, R, f = check(k, True)
            if ok:
                k_star, res = k, (R, f)
                break
        nonet = []
        for k in sorted({1, 5, 10, 20, args.max_cont} | ({k_star} if k_star else set())):
            ok, R, f = check(k, False)
            nonet.append(dict(cont=k, meets_target=ok, resid_max_over_target=round(float((R / R_plain).max()), 4),
                              frc=round(f, 4)))
        log("no-network continuation:", json.dumps(nonet))
        if k_star is None:
            log("learned start missed the target within %d continuation iterations: plain ICON is the result" % args.max_cont)
            learned = dict(k=None, fallback="all slices (plain ICON)", train_seconds=round(train_s, 1), no_network=nonet)
        else:
            dt_l = timed(lambda: solver.run(S, d_sino, d_ml, k_star), "learned")
            value = S * args.steps / dt_l
            ms_step = dt_l / args.steps * 1e3
            # verify the timed result and count slices that would fall back to plain ICON
            d_r = DeviceArray((S, k_star + 1), np.float64)
            solver.run(S, d_sino, d_ml, k_star, d_resid=d_r, d_reproj=d_rp)
            L.icon_device_synchronize()
            R = d_r.numpy()[:, -1]
            _, f = warmstart.reprojection_frc(d_rp.numpy(), cfg["sino"], iomit, band)
            fallback = int((R > R_plain).sum())
            learned = dict(start_iterations=args.start, continuation_iterations=k_star,
                           slices_per_s=round(value, 3), speedup_vs_plain=round(value / plain_value, 3),
                           resid_mean=float(R.mean()), resid_max_over_target=round(float((R / R_plain).max()), 5),
                           frc=round(f, 5), frc_plain=round(frc_plain, 5), frc_band=list(band),
                           fallback_slices=fallback, train_steps=args.train_steps,
                           train_loss_last=round(float(np.mean(losses[-50:])), 4), train_seconds=round(train_s, 1),
                           unet_batch=args.unet_batch, no_network=nonet)
            log("learned:", json.dumps(learned))
            # the U-Net forward alone (the largest part of a learned step)
            dt_u = timed(lambda: solver.net.warm_start_dev(S, solver.d_start, solver.d_init), "unet forward")
            gf = unet_gflop_per_slice()
            tf = gf * S * args.steps / dt_u / 1e3
            unet_roofline = dict(bound="mfma", achieved=round(tf, 1), peak=BF16_DENSE_PEAK_TFLOPS, unit="TFLOP/s",
                                 frac=round(tf / BF16_DENSE_PEAK_TFLOPS, 4),
                                 kernel="U-Net forward, all layers (BF16 MFMA convolutions), batches of %d" % args.unet_batch,
                                 ms_per_256_slices=round(dt_u / args.steps * 1e3 * 256 / S, 2),
                                 gflop_per_slice=round(gf, 2))
            log("unet forward:", json.dumps(unet_roofline))
        solver.close()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: %d slices x %d iterations" % (args.cpu_slices, args.cpu_iters))
        cpu = cpu_baseline(cfg, args.cpu_slices, args.cpu_iters, iters, min(args.cpu_threads, os.cpu_count() or 1))
        log("cpu baseline done")
    workload = ("learned-start ICON vs plain ICON: %d slices of 512x512, 41 tilts +-60 deg step 3, SNR 0.05; "
                "target = per-slice data residual of plain ICON (%d iterations %d,%d,%d, m >= 0) with reprojection "
                "FRC within %.3f; sinograms resident in HBM, solves replayed as HIP graphs"
                % (S, iters, pre, icon_it, post, args.frc_tol))
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
        "data": "synthetic (seeded blob phantoms, analytic projections, Gaussian noise)",
        "config": {
            "workload": workload,
            "slices": S, "N": 512, "tilts": 41, "plain_iterations": iters,
            "plain_slices_per_s": round(plain_value, 3),
            "plain_resid_mean": float(R_plain.mean()), "plain_resid_max": float(R_plain.max()),
            "plain_frc": round(frc_plain, 5),
            "learned": learned,
            "col_block": int(plan.info.col_block),
            "autotune": autotune,
            "ms_per_iteration_kernels": round(per_iter_ms, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "iteration_roofline": iteration_roofline(plan, S, per_iter_ms),
        "unet_roofline": unet_roofline,
        "kb6_variant": kb6,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
