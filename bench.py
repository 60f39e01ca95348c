#!/usr/bin/env python3
"""Throughput of the MI355X ICON core on the batched 512^2 tilt-series workload:
slices/s to a fixed data residual, learned start against plain ICON.

Workload: every slice of a synthetic tilt series (256 slices of 512 x 512,
41 tilts +-60 deg in 3 deg steps, SNR 0.05), sinograms resident in HBM.

* Plain leg: one step = the batched plain ICON solve (200 iterations split
  20 INFR / 160 ICON / 20 INFR, threshold m >= 0) replayed as one HIP graph,
  with its per-slice data residual R(m) = sqrt(sum w|Am-b|^2 / sum w|b|^2)
  (FP64) and reprojection, the residuals read back.  Each slice's final R is
  the target.
* Learned leg: a U-Net trained on the device (its own BF16 MFMA kernels) on
  (10-iteration, 200-iteration) pairs of 20480 slices of other seeded
  phantoms; one step = 10 plain iterations, the U-Net warm start, k descent
  iterations with residuals and reprojection, the residuals read back, and
  plain ICON re-solving every slice whose R misses its target (the per-slice
  fallback, timed as part of the step).  k, with an output gain g of the
  warm start (the start is g (m + s U(m / s))), is calibrated once on a
  held-out phantom of the training distribution, never on the evaluated
  series: the smallest k at which, for some g of --gains, every calibration
  slice reaches 99 % of its plain-ICON residual (the margin keeps fallbacks
  on the evaluated series rare) with an FRC against the phantom no more than
  0.005 below plain ICON's; of those g, the one with the lowest worst-slice
  residual.
* Quality (not timed): per slice, relative L2 and FRC of the learned result
  against plain ICON's and against the phantom, and the same for the 10 + k
  iterations without the network.
* Short-start variant (short_start_variant, not the headline): the same
  learned leg with the network trained on, and started from, the
  2-iteration iterate (--short-start).

Prints one JSON line: value = learned-start slices/s, plus the plain leg,
the roofline of the dominant ICON kernel (algorithmic bytes / HIP-event-timed
launch duration against 8 TB/s HBM) and the CPU baseline (the FP32-mirror
oracle on the host cores, a fixed sample of the same workload solved in
full).
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
        # the update: image and Re(rhat) in, image out (f32); + the N0 non-zero grid rows out
        "row_fwd": 12 * N0 * N1 + 8 * N0 * n1,
        # N0 grid rows in + A^H W b in (c64) + Re(rhat) out (f32, all the update reads) + the N0 grid
        # rows of A rhat out
        "row_adj": 8 * N0 * n1 + 12 * N0 * N1 + 8 * N0 * n1,
        # interpolation partials in (<= 2 per node) + data + weighted values out
        "combine": 16 * M + 8 * M + 8 * M,
        # the step size: interpolation partials of A rhat in + the row sums of |rhat|^2 (f64)
        "combine_alpha": 16 * M + 8 * N0,
    }


# row_adj also does the row pass of A rhat (fused), so row_fwd runs once; the
# second combine is the step size's
LAUNCHES_PER_ITER = {"row_fwd": 1, "col_fwd": 2, "combine": 1, "col_adj": 1, "row_adj": 1, "combine_alpha": 1}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, sched, threads):
    """The FP32-mirror oracle (OpenMP, one slice per thread) on `threads` host
    threads: `threads` slices of the 512^2 series solved in full (200
    iterations, no extrapolation) -> slices/s in the metric's unit; and all
    of the 256^2 config (one slice, 61 tilts, 100 iterations) timed beside it."""
    import oracle
    pre, icon_it, post = sched
    plan = oracle.Plan(cfg["N"], cfg["N"], cfg["tilt"], 32)
    S_cpu = threads
    t0 = time.perf_counter()
    _, used = plan.solve_batch(cfg["sino"][:S_cpu], pre, icon_it, post, 2, 0.0, threads=threads)
    dt = time.perf_counter() - t0
    from feilab_amd import phantom
    c1 = phantom.config1()
    p1 = oracle.Plan(256, 256, c1["tilt"], 32)
    t1 = time.perf_counter()
    p1.solve(c1["sino"], c1["pre"], c1["icon"], c1["post"], 2, 0.0)
    dt1 = time.perf_counter() - t1
    return dict(value=round(S_cpu / dt, 4), unit="slices/s", cores=int(used), kind="port",
                cpu_model=cpu_model(), host_cpus=os.cpu_count(),
                sample="%d slices of the same 512^2 x %d-tilt series, all %d iterations (%d,%d,%d), FP32-mirror "
                       "oracle, OpenMP one slice per thread on %d threads; %.1f s wall"
                       % (S_cpu, len(cfg["tilt"]), pre + icon_it + post, pre, icon_it, post, used, dt),
                config1_seconds=round(dt1, 2),
                config1="one 256^2 slice, 61 tilts, 100 iterations (10,80,10), one thread")


COPY_BYTES = 1 << 30


def iteration_roofline(plan, S, per_iter_ms):
    """One ICON iteration (7 launches) against HBM with this build's own
    algorithmic bytes per slice: the per-kernel bytes of algorithmic_bytes()
    times the launches per iteration.  The reference's dataflow (17 G + 6 S +
    8 I_c + 3 I_r per slice: three FFTs, pads, spread, two interpolations,
    weights, residual and update as separate passes) is reported beside it
    as an equivalent rate, not as a fraction of the peak."""
    kb = algorithmic_bytes(plan)
    own = sum(kb[k] * n for k, n in LAUNCHES_PER_ITER.items())
    ach = own * S / (per_iter_ms * 1e-3) / 1e9
    G = 8 * plan.n0 * plan.n1
    ref = 17 * G + 6 * 8 * plan.M + 8 * 8 * plan.N0 * plan.N1 + 3 * 4 * plan.N0 * plan.N1
    return dict(bytes_per_slice=own, ms_per_iteration=round(per_iter_ms, 4), achieved=round(ach, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                reference_dataflow_bytes_per_slice=ref,
                reference_dataflow_equivalent_GBs=round(ref * S / (per_iter_ms * 1e-3) / 1e9, 1))


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


def counter_pass(counters, target, S, kernel_substr):
    """Average per-dispatch values of TCC counters (summed over all instances)
    for the dispatches whose name contains kernel_substr, from one
    rocprofv3 --pmc child run of the probe mode; None if the profiler is
    unavailable or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="bench_pmc_")
    cmd = [exe, "--pmc"] + list(counters) + ["--output-format", "csv", "-d", d, "-o", "p", "--",
                                             sys.executable, os.path.abspath(__file__), "--probe", target,
                                             "--slices", str(S)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120,
                           env=dict(os.environ, TMPDIR=d))
    except subprocess.TimeoutExpired:
        log("counter pass %s on %s timed out" % (counters, target))
        return None
    vals = {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for row in csv.DictReader(open(os.path.join(root, f))):
                    if kernel_substr in row["Kernel_Name"] and row["Counter_Name"] in counters:
                        key = (row["Counter_Name"], row["Dispatch_Id"])
                        vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    shutil.rmtree(d, ignore_errors=True)
    if r.returncode not in (0, 139) and not vals:   # 139: the profiler's own exit after complete results
        log("counter pass %s on %s failed (rc %d)" % (counters, target, r.returncode))
        return None
    out = {}
    for c in counters:
        v = [x for (n, _), x in vals.items() if n == c]
        if not v:
            return None
        out[c] = float(np.mean(v))
    return out


READ_COUNTERS = ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_BUBBLE_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def measure_traffic(kernel, S):
    """HBM bytes per launch of `kernel` from the L2-to-memory request counters
    split by size: reads = 128 TCC_BUBBLE + u (RDREQ - 32B - BUBBLE) + 32 RDREQ_32B,
    writes = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B), where u, the bytes of a
    plain read request, is calibrated on a 1 GiB copy (its reads then equal
    1 GiB; on gfx950 u = 128, a full L2 line) and the copy's write bytes are
    checked against 1 GiB the same way."""
    name = {"col_fwd": "k_col_fwd", "col_adj": "k_col_adj", "row_fwd": "k_row_fwd", "row_adj": "k_row_adj"}[kernel]
    rk = counter_pass(READ_COUNTERS, kernel, S, name)
    wk = counter_pass(WRITE_COUNTERS, kernel, S, name)
    rc = counter_pass(READ_COUNTERS, "copy", S, "k_copy_f4")
    wc = counter_pass(WRITE_COUNTERS, "copy", S, "k_copy_f4")
    if None in (rk, wk, rc, wc):
        return None, None
    plain_c = rc["TCC_EA0_RDREQ_sum"] - rc["TCC_EA0_RDREQ_32B_sum"] - rc["TCC_BUBBLE_sum"]
    if plain_c <= 0:
        return None, None
    u = (COPY_BYTES - 128 * rc["TCC_BUBBLE_sum"] - 32 * rc["TCC_EA0_RDREQ_32B_sum"]) / plain_c
    wbytes = lambda w: 64 * w["TCC_EA0_WRREQ_64B_sum"] + 32 * (w["TCC_EA0_WRREQ_sum"] - w["TCC_EA0_WRREQ_64B_sum"])
    rbytes = lambda r: (128 * r["TCC_BUBBLE_sum"] + u * (r["TCC_EA0_RDREQ_sum"] - r["TCC_EA0_RDREQ_32B_sum"] -
                                                        r["TCC_BUBBLE_sum"]) + 32 * r["TCC_EA0_RDREQ_32B_sum"])
    rd, wr = rbytes(rk), wbytes(wk)
    detail = dict(read_bytes=int(rd), write_bytes=int(wr), read_request_bytes=round(u, 2),
                  copy_write_over_1GiB=round(wbytes(wc) / COPY_BYTES, 4),
                  counters={k: round(v, 1) for k, v in list(rk.items()) + list(wk.items())},
                  copy_counters={k: round(v, 1) for k, v in list(rc.items()) + list(wc.items())})
    return rd + wr, detail


def pair_chunks(total, chunk, seed0, workers):
    """Sinograms of seeded training phantoms (config-2 series of `chunk`
    slices, seeds seed0, seed0 + 1, ...), generated ahead of use by a pool of
    fresh interpreter processes."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    from feilab_amd import phantom
    jobs = [(min(chunk, total - k), seed0 + i) for i, k in enumerate(range(0, total, chunk))]
    with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn")) as ex:
        yield from ex.map(phantom.config2_sino, jobs)


def band_mean(curves, band):
    return np.asarray(curves)[:, band[0]:band[1]].mean(axis=1)


def rel_rows(a, b):
    a = np.asarray(a, np.float32).reshape(len(a), -1)
    b = np.asarray(b, np.float32).reshape(len(b), -1)
    return np.sqrt(np.einsum("ij,ij->i", a - b, a - b, dtype=np.float64) / np.einsum("ij,ij->i", b, b,
                                                                                        dtype=np.float64))


def learned_leg(args, plan, cfg, d_sino, sched, R_plain, m_plain, frc_plain, plain_value, timed, start):
    """Trains the warm start on (start, final)-iteration pairs, calibrates k
    on a held-out phantom, times the learned solve with its fallback on the
    evaluated series and measures its quality.  Returns (learned dict,
    unet_roofline dict, (value, ms_step) or None)."""
    import feilab_amd
    from feilab_amd import frc as F
    from feilab_amd import phantom, warmstart
    from feilab_amd.icon import DeviceArray, SolveOptions
    L = feilab_amd.lib()
    S = d_sino.shape[0]
    N = plan.N0
    band = (1, N // 4)
    iters = sum(sched)
    workers = max(1, min(8, (os.cpu_count() or 2) - 1))

    # training pairs and the network
    log("training pairs: %d slices of seeded phantoms, (%d, %d) iterations, %d generator processes"
        % (args.train_pairs, start, iters, workers))
    t0 = time.perf_counter()
    x, t = warmstart.make_pairs(plan, pair_chunks(args.train_pairs, args.pair_chunk, 5001, workers),
                                args.train_pairs, start, sched, log=log)
    pairs_s = time.perf_counter() - t0
    log("pairs ready in %.1f s" % pairs_s)
    t0 = time.perf_counter()
    params, losses = warmstart.train_on_pairs(x, t, steps=args.train_steps, log=log)
    train_s = time.perf_counter() - t0
    del x, t
    solver = warmstart.LearnedSolver(plan, params, S, start=start, schedule=sched, batch=args.unet_batch,
                                     overlap=bool(args.overlap))

    # continuation length (and output gain) from a held-out phantom of the training distribution
    Sc = args.calib_slices
    cal = phantom.config2(nslices=Sc, seed=3001)
    truth_c = phantom.render_slices_3d(cal["blobs"], N, N, np.arange(Sc))
    d_c = DeviceArray.from_host(cal["sino"])
    d_mc = DeviceArray((Sc, N, N), np.float32)
    d_rc = DeviceArray((Sc, max(iters, args.max_cont) + 1), np.float64)
    plan.reconstruct_dev(Sc, d_c, d_mc, SolveOptions(*sched, 2, 0.0, 1, 0), d_resid=d_rc)
    L.icon_device_synchronize()
    R_c = d_rc.numpy().reshape(-1)[:Sc * (iters + 1)].reshape(Sc, iters + 1)[:, -1]
    q_plain_c = band_mean(F.frc_slices(d_mc.numpy(), truth_c), band)
    # the smallest continuation k at which, for some output gain g of the
    # network's start (icon_unet_set_output_gain), every calibration slice
    # meets the residual and FRC targets; of the gains that do, the one with
    # the lowest worst-slice residual
    gains = [float(v) for v in args.gains.split(",")]
    k_star, g_star, scan = None, None, []
    for k in range(1, args.max_cont + 1):
        passing = []
        for g in gains:
            solver.set_gain(g)
            solver.run(Sc, d_c, d_mc, k, d_resid=d_rc)
            L.icon_device_synchronize()
            Rk = d_rc.numpy().reshape(-1)[:Sc * (k + 1)].reshape(Sc, k + 1)[:, -1]
            qk = band_mean(F.frc_slices(d_mc.numpy(), truth_c), band)
            row = dict(k=k, gain=g, resid_max_over_target=round(float((Rk / R_c).max()), 4),
                       frc_truth_drop_max=round(float((q_plain_c - qk).max()), 4),
                       frc_truth_mean=round(float(qk.mean()), 4))
            scan.append(row)
            log("calibration k=%d gain=%.2f:" % (k, g), json.dumps(row))
            if (Rk <= (1.0 - args.calib_margin) * R_c).all() and (qk >= q_plain_c - args.frc_tol).all():
                passing.append((float((Rk / R_c).max()), g))
        if passing:
            k_star, g_star = k, min(passing)[1]
            break
    if g_star is not None:
        solver.set_gain(g_star)
    calib = dict(phantom_seed=3001, slices=Sc, resid_margin=args.calib_margin, gains=gains,
                 frc_truth_plain_mean=round(float(q_plain_c.mean()), 4), scan=scan)
    for a in (d_c, d_mc, d_rc):
        a.free()
    if k_star is None:
        log("no continuation up to %d meets the calibration target: plain ICON is the result" % args.max_cont)
        solver.close()
        return dict(continuation_iterations=None, fallback="all slices (plain ICON)", calibration=calib,
                    train_seconds=round(train_s, 1), pairs_seconds=round(pairs_s, 1)), None, None

    # the timed learned solve on the evaluated series, fallback included
    d_ml = DeviceArray((S, N, N), np.float32)
    d_rl = DeviceArray((S, k_star + 1), np.float64)
    d_rpl = DeviceArray((S, N), np.float32)
    fallbacks = []

    def learned_step():
        fallbacks.append(len(solver.solve(S, d_sino, d_ml, k_star, R_plain, d_rl, d_rpl)))

    dt_l = timed(learned_step, "learned")
    value = S * args.steps / dt_l
    ms_step = dt_l / args.steps * 1e3
    R_l = d_rl.numpy()[:, -1]
    m_l = d_ml.numpy()
    serial = None
    if solver.overlap:
        # the same step with the two halves one after the other: same results, its time
        solver.overlap = False
        dt_s = timed(learned_step, "learned, halves not overlapped")
        same = bool(np.array_equal(d_ml.numpy(), m_l))
        solver.overlap = True
        serial = dict(slices_per_s=round(S * args.steps / dt_s, 3), ms_per_step=round(dt_s / args.steps * 1e3, 3),
                      same_result=same)
        log("serial learned step:", json.dumps(serial))
    _, frc_l = warmstart.reprojection_frc(d_rpl.numpy(), cfg["sino"], plan.iomit, band)

    # quality: against plain ICON and against the phantom; the same iterations without the network
    truth = phantom.render_slices_3d(cfg["blobs"], N, N, np.arange(S))
    solver.run(S, d_sino, d_ml, k_star, d_resid=d_rl, d_reproj=d_rpl, use_net=False)
    L.icon_device_synchronize()
    m_n, R_n = d_ml.numpy(), d_rl.numpy()[:, -1]
    _, frc_n = warmstart.reprojection_frc(d_rpl.numpy(), cfg["sino"], plan.iomit, band)
    q_p = band_mean(F.frc_slices(m_plain, truth), band)
    q_l = band_mean(F.frc_slices(m_l, truth), band)
    q_n = band_mean(F.frc_slices(m_n, truth), band)
    f_lp = band_mean(F.frc_slices(m_l, m_plain), band)
    f_np = band_mean(F.frc_slices(m_n, m_plain), band)
    e_lp, e_np = rel_rows(m_l, m_plain), rel_rows(m_n, m_plain)
    e_pt, e_lt, e_nt = rel_rows(m_plain, truth), rel_rows(m_l, truth), rel_rows(m_n, truth)
    st = lambda v: dict(mean=round(float(np.mean(v)), 4), min=round(float(np.min(v)), 4),
                        max=round(float(np.max(v)), 4))
    quality = dict(
        frc_vs_phantom=dict(plain=st(q_p), learned=st(q_l), no_network=st(q_n),
                            learned_minus_plain=st(q_l - q_p), no_network_minus_plain=st(q_n - q_p)),
        rel_l2_vs_phantom=dict(plain=st(e_pt), learned=st(e_lt), no_network=st(e_nt)),
        vs_plain_icon=dict(learned_rel_l2=st(e_lp), no_network_rel_l2=st(e_np), learned_frc=st(f_lp),
                           no_network_frc=st(f_np)),
        frc_band=list(band),
        note="per slice; FRC = mean over rings %d..%d of calFRC's ring correlation of the two slices" % band)
    learned = dict(start_iterations=start, continuation_iterations=k_star, output_gain=g_star,
                   slices_per_s=round(value, 3), speedup_vs_plain=round(value / plain_value, 3),
                   resid_mean=float(R_l.mean()), resid_max_over_target=round(float((R_l / R_plain).max()), 5),
                   fallback_slices_per_step=fallbacks[-1], fallback_slices_all_steps=fallbacks,
                   reprojection_frc=round(frc_l, 5), reprojection_frc_plain=round(frc_plain, 5),
                   no_network=dict(resid_max_over_target=round(float((R_n / R_plain).max()), 4),
                                   reprojection_frc=round(frc_n, 5)),
                   quality=quality, calibration=calib, train_pairs=args.train_pairs, train_steps=args.train_steps,
                   train_loss_first=round(float(np.mean(losses[:20])), 4),
                   train_loss_last=round(float(np.mean(losses[-50:])), 4),
                   pairs_seconds=round(pairs_s, 1), train_seconds=round(train_s, 1), unet_batch=args.unet_batch,
                   overlapped_halves=bool(solver.overlap), without_overlap=serial)
    log("learned:", json.dumps(learned))

    # the U-Net forward alone
    dt_u = timed(lambda: solver.net.warm_start_dev(S, solver.d_start, solver.d_init), "unet forward")
    gf = unet_gflop_per_slice()
    tf = gf * S * args.steps / dt_u / 1e3
    unet_roofline = dict(bound="mfma", achieved=round(tf, 1), peak=BF16_DENSE_PEAK_TFLOPS, unit="TFLOP/s",
                         frac=round(tf / BF16_DENSE_PEAK_TFLOPS, 4),
                         kernel="U-Net forward, all layers (BF16 MFMA convolutions), batches of %d" % args.unet_batch,
                         ms_per_256_slices=round(dt_u / args.steps * 1e3 * 256 / S, 2),
                         gflop_per_slice=round(gf, 2))
    log("unet forward:", json.dumps(unet_roofline))
    for a in (d_ml, d_rl, d_rpl):
        a.free()
    solver.close()
    return learned, unet_roofline, (value, ms_step)


def config3_bench(args):
    """The uneven batch (config 3): all 48 series of the generator in one
    packed solve against the same solves one series (one graph) at a time;
    prints one JSON line.  The slice counts can be scaled down (--c3-scale)
    to fit smaller devices; the packed and per-series results are compared."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    import feilab_amd
    from feilab_amd import phantom
    from feilab_amd.icon import DeviceArray, SolveOptions
    L = feilab_amd.lib()
    specs = phantom.config3_specs()
    for sp in specs:
        sp["S"] = max(1, int(round(sp["S"] * args.c3_scale)))
    pre, icon_it, post = [int(v) for v in args.c3_iters.split(",")]
    iters = pre + icon_it + post
    log("config 3: %d series, %d slices, widths %d..%d, %d iterations"
        % (len(specs), sum(sp["S"] for sp in specs), min(sp["N"] for sp in specs), max(sp["N"] for sp in specs), iters))
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=max(1, min(8, (os.cpu_count() or 2) - 1)),
                             mp_context=multiprocessing.get_context("spawn")) as ex:
        sinos = list(ex.map(phantom.config3_sino, specs))
    log("sinograms ready in %.1f s" % (time.perf_counter() - t0))
    plans = [feilab_amd.Plan(sp["N"], sp["tilt"], max_slices=sp["S"]) for sp in specs]
    dev_bytes = sum(p.info.device_bytes for p in plans)
    d_s = [DeviceArray.from_host(x) for x in sinos]
    d_m = [DeviceArray((sp["S"], sp["N"], sp["N"]), np.float32) for sp in specs]
    d_r = [DeviceArray((sp["S"], iters + 1), np.float64) for sp in specs]
    log("plans: %.1f GB of device memory" % (dev_bytes / 1e9))
    batch = feilab_amd.Batch(plans, [sp["S"] for sp in specs])
    cols = batch.autotune()
    opts = SolveOptions(pre, icon_it, post, 2, 0.0, 1, 0)
    total = sum(sp["S"] for sp in specs)

    def timed(step, what):
        for _ in range(args.warmup):
            step()
        L.icon_device_synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        L.icon_device_synchronize()
        dt = (time.perf_counter() - t) / args.steps
        log("%s: %.3f s per solve" % (what, dt))
        return dt

    dt_packed = timed(lambda: batch.reconstruct_dev(d_s, d_m, opts, d_resids=d_r), "packed")
    m_packed = [d.numpy() for d in d_m[:6]]
    r_packed = [d.numpy() for d in d_r]

    def per_series():
        for i, p in enumerate(plans):
            p.reconstruct_dev(specs[i]["S"], d_s[i], d_m[i], opts, d_resid=d_r[i])

    dt_series = timed(per_series, "one series at a time")
    same = all(np.array_equal(d.numpy(), m) for d, m in zip(d_m[:6], m_packed)) and \
        all(np.array_equal(d.numpy(), r) for d, r in zip(d_r, r_packed))
    classes = sorted({(p.n0, c) for p, c in zip(plans, cols)})
    out = dict(metric="config-3 slices/s, all series in packed launches vs one series at a time", unit="slices/s",
               value=round(total / dt_packed, 2), per_series_slices_per_s=round(total / dt_series, 2),
               speedup=round(dt_series / dt_packed, 3), same_result=bool(same), series=len(specs), slices=total,
               iterations=iters, schedule=[pre, icon_it, post], launches_per_iteration=batch.launches_per_iteration(),
               launches_per_iteration_one_at_a_time=7 * len(specs), grid_and_tile_classes=classes,
               device_gb=round(dev_bytes / 1e9, 1), slice_scale=args.c3_scale, steps=args.steps, warmup=args.warmup,
               resid_final_mean=float(np.mean([r[:, -1].mean() for r in r_packed])))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--iters", default="20,160,20", help="pre,icon,post")
    ap.add_argument("--start", type=int, default=10, help="plain iterations before the U-Net")
    ap.add_argument("--short-start", type=int, default=2,
                    help="start length of a second learned leg reported as short_start_variant (0: none)")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--train-pairs", type=int, default=20480, help="(start, final) training pairs, slices")
    ap.add_argument("--pair-chunk", type=int, default=256, help="slices per training phantom")
    ap.add_argument("--calib-slices", type=int, default=64, help="slices of the held-out calibration phantom")
    ap.add_argument("--unet-batch", type=int, default=128, help="slices per U-Net forward launch sequence")
    ap.add_argument("--overlap", type=int, default=0,
                    help="1: the learned solve in two halves on two streams (iterations overlap the U-Net)")
    ap.add_argument("--calib-margin", type=float, default=0.01,
                    help="calibration target: R <= (1 - margin) x plain ICON's R on every calibration slice")
    ap.add_argument("--max-cont", type=int, default=40, help="continuation iterations tried")
    ap.add_argument("--gains", default="1,0.95,0.9,0.85,0.8",
                    help="output gains of the warm start tried with each continuation length")
    ap.add_argument("--frc-tol", type=float, default=0.005,
                    help="allowed drop of the FRC against the phantom below plain ICON's")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads (and slices) of the CPU baseline; 0: OMP_NUM_THREADS or the CPU count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--col-block", type=int, default=0, help="grid columns per spread/interp tile (0: default)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-kb", action="store_true", help="skip the Kaiser-Bessel window variant")
    ap.add_argument("--probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--workload", default="config2", choices=("config2", "config3"),
                    help="config3: the uneven batch of 48 series, packed vs one series at a time")
    ap.add_argument("--c3-iters", default="5,10,5", help="config3: pre,icon,post")
    ap.add_argument("--c3-scale", type=float, default=1.0, help="config3: factor on every series' slice count")
    args = ap.parse_args()
    if args.probe:
        probe(args.probe, args.slices)
        return
    if args.workload == "config3":
        if args.gpus != 1:
            print("bench.py measures one GPU (--gpus %d is not supported)" % args.gpus)
            sys.exit(2)
        config3_bench(args)
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

    # ---- plain leg: the solve with residuals and reprojection, the residuals read back
    def plain_step():
        plan.reconstruct_dev(S, d_sino, d_m, opts, d_resid=d_res, d_reproj=d_rp)
        return d_res.numpy()

    dt_plain = timed(plain_step, "plain")
    plain_value = S * args.steps / dt_plain
    R_plain = d_res.numpy()[:, -1]
    m_plain = d_m.numpy()
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
        log("counter passes (L2-to-memory read and write requests by size) on %s and a 1 GiB copy" % dom)
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
        learned, unet_roofline, lv = learned_leg(args, plan, cfg, d_sino, (pre, icon_it, post), R_plain, m_plain,
                                                 frc_plain, plain_value, timed, args.start)
        if lv is not None:
            value, ms_step = lv
    # ---- the same learned leg from an earlier iterate (fewer plain iterations before the network)
    short = None
    if not args.no_learned and args.short_start and args.short_start != args.start:
        log("learned leg from the %d-iteration iterate" % args.short_start)
        ls, _, lsv = learned_leg(args, plan, cfg, d_sino, (pre, icon_it, post), R_plain, m_plain, frc_plain,
                                 plain_value, timed, args.short_start)
        short = dict(note="the learned leg with the U-Net trained on (%d, %d)-iteration pairs and started from "
                          "the %d-iteration iterate; everything else as the headline leg"
                          % (args.short_start, iters, args.short_start),
                     slices_per_s=None if lsv is None else round(lsv[0], 3),
                     ms_per_step=None if lsv is None else round(lsv[1], 3), learned=ls)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log("cpu baseline: %d slices x %d iterations on %d threads" % (threads, iters, threads))
        cpu = cpu_baseline(cfg, (pre, icon_it, post), threads)
        log("cpu baseline:", json.dumps(cpu))
    workload = ("learned-start ICON vs plain ICON: %d slices of 512x512, 41 tilts +-60 deg step 3, SNR 0.05; "
                "target = per-slice data residual of plain ICON (%d iterations %d,%d,%d, m >= 0), the continuation "
                "length and the warm start's output gain calibrated on a held-out training-distribution phantom "
                "where every slice reaches 99%% of its plain-ICON residual and keeps its FRC against the phantom "
                "within %.3f of plain ICON's; U-Net "
                "trained on %d (%d, %d)-iteration pairs; "
                "per-slice fallback to plain ICON timed in the step; sinograms resident in HBM, solves replayed as "
                "HIP graphs" % (S, iters, pre, icon_it, post, args.frc_tol, args.train_pairs, args.start, iters))
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
        "short_start_variant": short,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
