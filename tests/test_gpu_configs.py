"""The remaining workloads of the benchmark definition, as parity cases.

* config 1 (one 256^2 slice, 61 tilts, 100 iterations, Kaiser-Bessel width 6
  and oversampling 2): the KB plan and the reference's Gaussian plan against
  the FP64 oracle.
* config 3 (an uneven batch: tilt series of different widths, slice counts,
  tilt ranges, steps and SNR): every series against the FP32 mirror of the
  reference (the parity bar) and its error against the FP64 oracle stated.
* config 4 (precision study at config-2 shapes): FP32 device solves of 512^2
  slices, 41 tilts, 200 iterations, against the all-FP64 device solve
  (icon_reconstruct_fp64) and the FP64 oracle: relative L2 and
  reprojection-FRC difference per slice; the FP64 device solve itself against
  the FP64 oracle (rounding-level agreement).
* the column-tile autotuner keeps results identical to a fixed tile.
"""
import numpy as np
import pytest

from feilab_amd import phantom, warmstart

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_config1_kb6_and_gauss_vs_fp64(icon, oracle_mod):
    cfg = phantom.config1()
    it = (cfg["pre"], cfg["icon"], cfg["post"])
    m64 = oracle_mod.Plan(256, 256, cfg["tilt"], 64).solve(cfg["sino"], *it)
    errs = {}
    for w in ("kb6", "gauss"):
        plan = icon.Plan(256, cfg["tilt"], max_slices=1, window=w)
        errs[w] = rel(plan.solve(cfg["sino"], *it), m64)
        plan.close()
    print("config 1, 100 iterations, rel L2 vs FP64 oracle: kb6 %.2e, gauss %.2e" % (errs["kb6"], errs["gauss"]),
          flush=True)
    assert errs["gauss"] < 1e-4 and errs["kb6"] < 2e-4


# (width, slices, max tilt, step, SNR)
UNEVEN = [(128, 3, 45, 1, 0.02), (200, 2, 70, 4, 0.2), (256, 2, 60, 2, 0.1), (96, 4, 55, 3, 0.05),
          (160, 3, 65, 2, 0.15), (300, 2, 50, 4, 0.08)]


def test_config3_uneven_batch(icon, oracle_mod):
    """Each series solved as one batched launch of its slices; per series the
    device result against the FP32 mirror (<= 1e-4) and the FP64 oracle."""
    rng = np.random.default_rng(3003)
    it = (5, 20, 5)
    for N, S, mt, st, snr in UNEVEN:
        tilt = phantom.tilt_series(mt, st)
        b = phantom.blobs_3d(N, S, N, 20, rng)
        sinos = phantom.add_noise(phantom.project_3d(b, N, np.arange(S), tilt), snr, rng)
        plan = icon.Plan(N, tilt, max_slices=S)
        m = plan.solve(sinos, *it)
        plan.close()
        o32, o64 = oracle_mod.Plan(N, N, tilt, 32), oracle_mod.Plan(N, N, tilt, 64)
        e32 = max(rel(m[s], o32.solve(sinos[s], *it)) for s in range(S))
        e64 = max(rel(m[s], o64.solve(sinos[s], *it)) for s in range(S))
        print("series %d x %d slices, %d tilts to %d deg: rel L2 vs FP32 mirror %.2e, vs FP64 %.2e"
              % (N, S, len(tilt), mt, e32, e64), flush=True)
        assert e32 < 1e-4 and e64 < 1e-4


def test_fp64_device_solve_vs_fp64_oracle(icon, oracle_mod):
    """The all-FP64 device solve restates the FP64 oracle's algorithm: images,
    residual and step-size traces and reprojections agree to rounding."""
    from feilab_amd.icon import reconstruct_fp64
    rng = np.random.default_rng(6006)
    N, S = 96, 2
    tilt = phantom.tilt_series(60, 3)
    sinos = phantom.add_noise(phantom.project_3d(phantom.blobs_3d(N, S, N, 12, rng), N, np.arange(S), tilt), 0.5, rng)
    it = (5, 20, 5)
    m, rp, res, al = reconstruct_fp64(N, tilt, sinos, *it)
    orc = oracle_mod.Plan(N, N, tilt, 64)
    for s in range(S):
        mo, reso, alo = orc.solve(sinos[s], *it, trace=True)
        e = rel(m[s], mo)
        print("FP64 device vs FP64 oracle, slice %d: rel L2 %.2e, R %.3e, alpha %.3e" % (
            s, e, abs(res[s][-1] - reso[-1]) / reso[-1], rel(al[s], alo)), flush=True)
        # the oracle hands its FP64 image back rounded to float32: that rounding
        # (about 3e-8 relative) bounds the image comparison, not the solve
        assert e < 1e-7
        assert np.allclose(res[s], reso, rtol=1e-9, atol=1e-14)
        assert rel(al[s], alo) < 1e-9
        assert rel(rp[s], orc.reproject(mo.astype(np.float32))) < 1e-6   # the oracle reprojects the float image


def test_config4_precision_study(icon, oracle_mod):
    from feilab_amd.icon import reconstruct_fp64
    S = 2
    cfg = phantom.config2(nslices=S)
    it = (cfg["pre"], cfg["icon"], cfg["post"])
    plan = icon.Plan(512, cfg["tilt"], max_slices=S)
    m, rp, res, _ = plan.solve(cfg["sino"], *it, trace=True, reproject=True)
    m64d, rp64d, res64d, _ = reconstruct_fp64(512, cfg["tilt"], cfg["sino"], *it)
    orc = oracle_mod.Plan(512, 512, cfg["tilt"], 64)
    m64, _ = orc.solve_batch(cfg["sino"], *it, threads=S)
    band = (1, 128)
    for s in range(S):
        _, f = warmstart.reprojection_frc(rp[s:s + 1], cfg["sino"][s:s + 1], plan.iomit, band)
        _, f64 = warmstart.reprojection_frc(rp64d[s:s + 1], cfg["sino"][s:s + 1], plan.iomit, band)
        e, e_orc, e_dev = rel(m[s], m64d[s]), rel(m[s], m64[s]), rel(m64d[s], m64[s])
        print("512^2 slice %d, 200 iterations: FP32 vs all-FP64 device rel L2 %.2e (R %.2e), vs FP64 oracle %.2e; "
              "FP64 device vs FP64 oracle %.2e; reprojection FRC %.5f vs %.5f"
              % (s, e, abs(res[s][-1] - res64d[s][-1]) / res64d[s][-1], e_orc, e_dev, f, f64), flush=True)
        # (FP64 device vs the oracle's float32-rounded FP64 image: rounding level)
        assert e < 1e-4 and e_orc < 1e-4 and e_dev < 1e-7 and abs(f - f64) < 1e-4


def test_autotuner_keeps_results(icon, oracle_mod):
    cfg = phantom.config1()
    plan = icon.Plan(256, cfg["tilt"], max_slices=4)
    it = (cfg["pre"], cfg["icon"], cfg["post"])
    sinos = np.stack([cfg["sino"]] * 4)
    m8 = plan.solve(sinos, *it)
    c, ms = plan.autotune(4)
    print("autotune: C = %d, ms per iteration's column kernels %s" % (c, ms), flush=True)
    assert c in (8, 16) and set(ms) == {8, 16} and ms[c] == min(ms.values())
    assert plan.info.col_block == c
    assert rel(plan.solve(sinos, *it), m8) < 1e-5
