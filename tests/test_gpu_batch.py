"""The packed solve of uneven tilt series (icon_batch_*), the benchmark's
config 3 at bounded slice counts.

* Every series of a packed solve equals its own single-plan solve bit for bit
  (same kernels, same per-slice arithmetic; only the launch packing differs),
  images, residual traces, step sizes and reprojections.
* Series of the config-3 generator (widths 160..1007, so grids of 512, 1024
  and 2048, tilt ranges 45..70 deg in steps of 1..4 deg, SNR 0.02..0.2) in
  one packed solve, each against the FP32 mirror of the reference (<= 1e-4
  relative L2, the parity bar) with its error against the FP64 oracle
  stated.
* The batch autotuner picks a column tile per size class and leaves the
  results unchanged to rounding.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from feilab_amd import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


# (width, thickness, slices, max tilt, step)
MIXED = [(160, 160, 2, 60, 3), (300, 200, 3, 45, 2), (520, 520, 2, 70, 4), (96, 96, 4, 55, 3), (700, 650, 1, 50, 4),
         (256, 256, 2, 60, 2)]


def test_packed_solve_equals_single_plans(icon):
    from feilab_amd.icon import DeviceArray, SolveOptions
    print("packed solve of %d series vs their single-plan solves" % len(MIXED), flush=True)
    rng = np.random.default_rng(4004)
    plans, sinos = [], []
    for N0, N1, S, mt, st in MIXED:
        tilt = phantom.tilt_series(mt, st)
        b = phantom.blobs_3d(N0, S, N1, 15, rng)
        sinos.append(phantom.add_noise(phantom.project_3d(b, N0, np.arange(S), tilt), 0.3, rng))
        plans.append(icon.Plan(N0, tilt, thickness=N1, max_slices=S + 1))
    batch = icon.Batch(plans, [m[2] for m in MIXED])
    print("launches per iteration: %d (single plans: %d)" % (batch.launches_per_iteration(), 7 * len(MIXED)),
          flush=True)
    it = (2, 4, 2)
    iters = sum(it)
    d_s = [DeviceArray.from_host(x) for x in sinos]
    d_m = [DeviceArray((m[2], p.N0, p.N1), np.float32) for m, p in zip(MIXED, plans)]
    d_r = [DeviceArray((m[2], iters + 1), np.float64) for m in MIXED]
    d_a = [DeviceArray((m[2], iters), np.float64) for m in MIXED]
    d_p = [DeviceArray((m[2], p.N0), np.float32) for m, p in zip(MIXED, plans)]
    opts = SolveOptions(*it, 2, 0.0, 1, 0)
    for rep in range(2):
        batch.reconstruct_dev(d_s, d_m, opts, d_reprojs=d_p, d_resids=d_r, d_alphas=d_a)
        icon.lib().icon_device_synchronize()
        for i, p in enumerate(plans):
            S = MIXED[i][2]
            e_m = DeviceArray((S, p.N0, p.N1), np.float32)
            e_r = DeviceArray((S, iters + 1), np.float64)
            e_a = DeviceArray((S, iters), np.float64)
            e_p = DeviceArray((S, p.N0), np.float32)
            p.reconstruct_dev(S, d_s[i], e_m, opts, d_reproj=e_p, d_resid=e_r, d_alpha=e_a)
            icon.lib().icon_device_synchronize()
            np.testing.assert_array_equal(d_m[i].numpy(), e_m.numpy())
            np.testing.assert_array_equal(d_r[i].numpy(), e_r.numpy())
            np.testing.assert_array_equal(d_a[i].numpy(), e_a.numpy())
            np.testing.assert_array_equal(d_p[i].numpy(), e_p.numpy())
        print("  replay %d: every series bitwise equal" % rep, flush=True)
    batch.close()


def test_config3_series_packed_vs_oracle(icon, oracle_mod):
    specs = phantom.config3_specs()
    # one series per grid class and the extremes of width and tilt count
    pick = sorted({min(range(48), key=lambda i: specs[i]["N"]), max(range(48), key=lambda i: specs[i]["N"]),
                   max(range(48), key=lambda i: len(specs[i]["tilt"])),
                   next(i for i in range(48) if 256 < specs[i]["N"] <= 512),
                   next(i for i in range(48) if 512 < specs[i]["N"] <= 700)})
    S = 2
    it = (3, 6, 3)
    sel = [specs[i] for i in pick]
    print("config-3 series %s: widths %s, tilts %s, %d slices each, packed"
          % (pick, [s["N"] for s in sel], [len(s["tilt"]) for s in sel], S), flush=True)
    sinos = [phantom.config3_sino(s, S) for s in sel]
    plans = [icon.Plan(s["N"], s["tilt"], max_slices=S) for s in sel]
    batch = icon.Batch(plans, [S] * len(sel))
    m, res, _ = batch.solve(sinos, *it, trace=True)
    batch.close()

    def oracle_errors(k):
        s = sel[k]
        o32 = oracle_mod.Plan(s["N"], s["N"], s["tilt"], 32)
        o64 = oracle_mod.Plan(s["N"], s["N"], s["tilt"], 64)
        m32, r32, _ = o32.solve(sinos[k][0], *it, trace=True)
        m64 = o64.solve(sinos[k][0], *it)
        return rel(m[k][0], m32), rel(m[k][0], m64), abs(res[k][0][-1] - r32[-1]) / r32[-1]

    with ThreadPoolExecutor(len(sel)) as ex:
        errs = list(ex.map(oracle_errors, range(len(sel))))
    for s, (e32, e64, er) in zip(sel, errs):
        print("  series %dx%d, %d tilts to %.1f deg, SNR %.3f: rel L2 vs FP32 mirror %.2e, vs FP64 %.2e, R %.1e"
              % (s["N"], s["N"], len(s["tilt"]), s["max_tilt"], s["snr"], e32, e64, er), flush=True)
        assert e32 < 1e-4 and e64 < 1e-4 and er < 1e-4


def test_batch_autotune_keeps_results(icon):
    rng = np.random.default_rng(5005)
    cfg = [(200, 2, 60, 3), (512, 2, 60, 3), (900, 1, 55, 4)]
    plans, sinos = [], []
    for N, S, mt, st in cfg:
        tilt = phantom.tilt_series(mt, st)
        sinos.append(phantom.add_noise(phantom.project_3d(phantom.blobs_3d(N, S, N, 15, rng), N, np.arange(S), tilt),
                                       0.3, rng))
        plans.append(icon.Plan(N, tilt, max_slices=S))
    batch = icon.Batch(plans, [c[1] for c in cfg])
    m8, _, _ = batch.solve(sinos, 2, 4, 2)
    cs = batch.autotune()
    print("batch autotune: column tiles %s" % cs, flush=True)
    assert all(c in (8, 16) for c in cs)
    m_t, _, _ = batch.solve(sinos, 2, 4, 2)
    for a, b in zip(m_t, m8):
        assert rel(a, b) < 1e-5
    batch.close()
