"""GPU parity of the HIP path against the oracle (FP32 mirror and FP64).

Every test calls the library through the C ABI (ctypes) and compares with the
CPU restatement of the reference on the same seeded inputs.  Tolerances:

* FFT kernels vs numpy (FP64): relative L2 <= 2e-6 (FP32 radix-16 Stockham).
* NUFFT forward / adjoint / 1D transform vs the FP32-mirror oracle: relative
  L2 <= 2e-5 (FFT rounding order differs from the oracle's radix-2).
* Whole ICON solves vs the FP32-mirror oracle: relative L2 <= 1e-4 per slice
  (the bar for "the reference's own reconstruction"), data residual within
  1e-4 relative.
"""
import os

import numpy as np
import pytest

from feilab_amd import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n", [128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_rows_vs_numpy(icon, n, sign):
    _print("fft_rows", n, sign)
    rng = np.random.default_rng(n + sign)
    x = (rng.standard_normal((37, n)) + 1j * rng.standard_normal((37, n))).astype(np.complex64)
    y = icon.fft_rows(x, sign)
    ref = np.fft.fft(x.astype(np.complex128), axis=-1) if sign < 0 else np.fft.ifft(x.astype(np.complex128), axis=-1) * n
    assert rel(y, ref) < 2e-6


@pytest.mark.parametrize("shape", [(256, 256), (512, 1024), (1024, 1024), (2048, 512)])
def test_fft_2d_vs_numpy(icon, shape):
    _print("fft_2d", shape)
    rng = np.random.default_rng(shape[0])
    x = (rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)).astype(np.complex64)
    y = icon.fft_2d(x, -1)
    ref = np.fft.fft2(x.astype(np.complex128))
    assert rel(y, ref) < 3e-6
    z = icon.fft_2d(x, 1)
    assert rel(z, np.fft.ifft2(x.astype(np.complex128)) * shape[0] * shape[1]) < 3e-6


GEOMS = [
    (64, 64, phantom.tilt_series(60, 6)),      # small, sigma = 2
    (100, 72, phantom.tilt_series(50, 5)),     # non-power-of-two width and thickness
    (256, 256, phantom.tilt_series(60, 2)),    # config 1 shape
]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "%dx%d_nz%d" % (g[0], g[1], len(g[2])))
def test_nufft_forward_adjoint_vs_oracle(icon, oracle_mod, geom):
    N0, N1, tilt = geom
    _print("nufft", N0, N1, len(tilt))
    rng = np.random.default_rng(N0 * 7 + N1)
    plan = icon.Plan(N0, tilt, thickness=N1, max_slices=3)
    orc = oracle_mod.Plan(N0, N1, tilt, 32)
    assert (plan.n0, plan.n1, plan.M, plan.iomit) == (orc.n0, orc.n1, orc.M, orc.iomit)
    img = rng.standard_normal((3, N0, N1)).astype(np.float32)
    f = plan.forward(img)
    for s in range(3):
        assert rel(f[s], orc.forward(img[s].astype(np.complex128))) < 2e-5
    cimg = (rng.standard_normal((2, N0, N1)) + 1j * rng.standard_normal((2, N0, N1))).astype(np.complex64)
    fc = plan.forward(cimg)
    for s in range(2):
        assert rel(fc[s], orc.forward(cimg[s])) < 2e-5
    g = (rng.standard_normal((3, plan.M)) + 1j * rng.standard_normal((3, plan.M))).astype(np.complex64)
    a = plan.adjoint(g)
    for s in range(3):
        assert rel(a[s], orc.adjoint(g[s])) < 2e-5
    # adjointness of the device operators themselves
    lhs = np.vdot(g[0].astype(np.complex128), fc[0])
    rhs = np.vdot(a[0].astype(np.complex128), cimg[0])
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("geom", GEOMS[:2], ids=lambda g: "%dx%d" % (g[0], g[1]))
def test_nufft_vs_fp64_ndft(icon, oracle_mod, geom):
    """The device NUFFT against the exact transform: the Gaussian-window error
    bound at m = 6, sigma >= 2 is 4 exp(-m pi (1 - 1/(2 sigma - 1))) ~ 1.4e-5."""
    N0, N1, tilt = geom
    rng = np.random.default_rng(3)
    plan = icon.Plan(N0, tilt, thickness=N1)
    orc = oracle_mod.Plan(N0, N1, tilt, 64)
    nodes = orc.tables()["nodes"]
    img = rng.standard_normal((N0, N1))
    exact = oracle_mod.ndft_forward(nodes, img.astype(np.complex128))
    assert rel(plan.forward(img.astype(np.float32)), exact) < 3e-5


def test_sino_to_b_vs_oracle(icon, oracle_mod):
    cfg = phantom.config1()
    plan = icon.Plan(256, cfg["tilt"], max_slices=2)
    orc = oracle_mod.Plan(256, 256, cfg["tilt"], 32)
    sinos = np.stack([cfg["sino"], cfg["clean"]])
    b = plan.sino_to_b(sinos)
    for s in range(2):
        assert rel(b[s], orc.sino_to_b(sinos[s])) < 1e-5


def _small_case(seed, N=96, nb=12, step=3):
    rng = np.random.default_rng(seed)
    tilt = phantom.tilt_series(60, step)
    b = phantom.blobs_2d(N, nb, rng)
    clean = phantom.project_2d(b, N, tilt)
    return tilt, phantom.add_noise(clean, 0.5, rng)


@pytest.mark.parametrize("data_type,thr", [(2, 0.0), (1, 0.03)])
def test_icon_solve_small_vs_oracle(icon, oracle_mod, data_type, thr):
    _print("solve small", data_type, thr)
    tilt, sino = _small_case(11)
    N = sino.shape[1]
    plan = icon.Plan(N, tilt, max_slices=1)
    orc = oracle_mod.Plan(N, N, tilt, 32)
    m, rp, res, al = plan.solve(sino, 5, 30, 5, data_type, thr, trace=True, reproject=True)
    mo, reso, alo = orc.solve(sino, 5, 30, 5, data_type, thr, trace=True)
    assert rel(m, mo) < 1e-4
    assert np.allclose(res, reso, rtol=1e-4, atol=1e-7)
    assert np.allclose(al, alo, rtol=1e-4)
    assert rel(rp, orc.reproject(mo)) < 1e-4


def test_icon_solve_config1_vs_oracle(icon, oracle_mod):
    """Config 1 (256^2, 61 tilts, 100 iterations) against the FP32 mirror."""
    cfg = phantom.config1()
    _print("solve config1")
    plan = icon.Plan(256, cfg["tilt"], max_slices=1)
    orc = oracle_mod.Plan(256, 256, cfg["tilt"], 32)
    m, res, al = plan.solve(cfg["sino"], cfg["pre"], cfg["icon"], cfg["post"], 2, 0.0, trace=True)
    mo, reso, alo = orc.solve(cfg["sino"], cfg["pre"], cfg["icon"], cfg["post"], 2, 0.0, trace=True)
    _print("config1 rel", rel(m, mo), "resid", res[-1], reso[-1])
    assert rel(m, mo) < 1e-4
    assert abs(res[-1] - reso[-1]) / reso[-1] < 1e-4


def test_bench_geometry_solve_vs_oracle(icon, oracle_mod):
    """The benchmark's geometry (512^2, 41 tilts +-60 step 3, a tilt range
    that leaves 14 of the 128 interpolation tiles without nodes) on two
    slices of the benchmark phantom, 20 iterations, against the FP32 mirror."""
    cfg = phantom.config2(nslices=2, nblobs=40)
    _print("solve bench geometry 512x512 x 41 tilts, 2 slices")
    plan = icon.Plan(512, cfg["tilt"], max_slices=2)
    orc = oracle_mod.Plan(512, 512, cfg["tilt"], 32)
    m, res, _ = plan.solve(cfg["sino"], 5, 10, 5, 2, 0.0, trace=True)
    for k in range(2):
        mo, reso, _ = orc.solve(cfg["sino"][k], 5, 10, 5, 2, 0.0, trace=True)
        _print("bench geometry slice", k, "rel", rel(m[k], mo), "resid", res[k][-1], reso[-1])
        assert rel(m[k], mo) < 1e-4
        assert abs(res[k][-1] - reso[-1]) / reso[-1] < 1e-4


def test_data_residual_vs_exact_transform(icon):
    """The device's R(m) (FP64 in k_combine) against numpy alone: the exact
    non-uniform DFT of the device's own iterate, the centred DFT of the
    projection rows and the density weights; R(0) = 1 exactly; on data that
    an image m* explains exactly R(m*) is the window error alone."""
    from exact_ref import centred_dft_rows, consistent_case, ndft_np, nodes_np
    _print("data residual vs exact transform")
    N = 64
    rng = np.random.default_rng(5)
    tilt = phantom.tilt_series(60, 6)
    sino = phantom.add_noise(phantom.project_2d(phantom.blobs_2d(N, 8, rng), N, tilt), 1.0, rng)
    plan = icon.Plan(N, tilt, max_slices=1)
    from feilab_amd.icon import geometry_tables
    w = geometry_tables(N, tilt)["w"].astype(np.float64)
    b = centred_dft_rows(sino)
    xy = nodes_np(N, tilt)
    for k in (0, 1, 4):
        m, res, _ = plan.solve(sino, k, 0, 0, 2, 0.0, trace=True)
        assert res[0] == 1.0
        R = np.sqrt(np.sum(w * np.abs(ndft_np(m, xy) - b) ** 2) / np.sum(w * np.abs(b) ** 2))
        _print("k", k, "device R", res[-1], "exact", R)
        assert abs(res[-1] - R) < 5e-6
    ms, p = consistent_case(N, tilt)
    _, res, _ = plan.solve(p, 0, 0, 0, 2, 0.0, m_init=ms, trace=True)
    assert res[0] < 1e-5


def test_icon_phase_bounds_on_device(icon):
    """The clamp runs in iterations pre <= it < iters - post (Iteration2.cu:129),
    bitwise: pre = 2, icon = 1 is the 3-step unclamped solve clamped once, and
    a following free step (post = 1) equals a warm-started 1-step solve."""
    tilt, s0 = _small_case(17, N=64, nb=8)
    plan = icon.Plan(64, tilt, max_slices=1)
    unclamped = plan.solve(s0, 3, 0, 0, 2, 0.0)
    assert (unclamped < 0).sum() > 50
    clamped = plan.solve(s0, 2, 1, 0, 2, 0.0)
    np.testing.assert_array_equal(clamped, np.maximum(unclamped, 0.0))
    np.testing.assert_array_equal(plan.solve(s0, 2, 1, 0, 1, 0.01), np.minimum(unclamped, np.float32(0.01)))
    full = plan.solve(s0, 2, 1, 1, 2, 0.0)
    np.testing.assert_array_equal(full, plan.solve(s0, 0, 0, 1, 2, 0.0, m_init=clamped))
    assert (full < 0).sum() > 50


def test_batch_matches_single(icon):
    """Slices of a batch are independent: solving them together gives each
    slice's single-slice result."""
    tilt, s0 = _small_case(21)
    _, s1 = _small_case(22)
    N = s0.shape[1]
    plan = icon.Plan(N, tilt, max_slices=2)
    mb = plan.solve(np.stack([s0, s1]), 3, 10, 3, 2, 0.0)
    m0 = plan.solve(s0, 3, 10, 3, 2, 0.0)
    m1 = plan.solve(s1, 3, 10, 3, 2, 0.0)
    assert rel(mb[0], m0) < 1e-6 and rel(mb[1], m1) < 1e-6


def test_graph_replay_matches_eager(icon):
    """The captured solve, replayed several times with eager work in between,
    gives the eager result every time."""
    tilt, s0 = _small_case(31)
    N = s0.shape[1]
    plan = icon.Plan(N, tilt, max_slices=1)
    eager = plan.solve(s0, 3, 10, 3, 2, 0.0, use_graph=False)
    from feilab_amd.icon import DeviceArray, SolveOptions
    d_s = DeviceArray.from_host(s0[None])
    d_m = DeviceArray((1, N, N), np.float32)
    opts = SolveOptions(3, 10, 3, 2, 0.0, 1)
    for rep in range(3):
        d_m.zero()
        plan.reconstruct_dev(1, d_s, d_m, opts)
        icon.lib().icon_device_synchronize()
        assert rel(d_m.numpy()[0], eager) < 1e-6, rep
        plan.solve(s0, 1, 1, 1, 2, 0.0)  # eager work between replays


def test_split_solve_with_reuse_data_matches_full(icon):
    """A solve stopped after k iterations and continued from its iterate, with
    the data set-up reused, equals the uninterrupted solve (the iteration
    state is m alone); three captured schedules alternate in the graph cache."""
    _print("split solve reuse_data")
    from feilab_amd.icon import DeviceArray, SolveOptions
    tilt, s0 = _small_case(29)
    _, s1 = _small_case(30)
    N = s0.shape[1]
    plan = icon.Plan(N, tilt, max_slices=2)
    sino = np.stack([s0, s1])
    full = plan.solve(sino, 4, 12, 4, 2, 0.0)
    d_s = DeviceArray.from_host(sino)
    d_a = DeviceArray((2, N, N), np.float32)
    d_b = DeviceArray((2, N, N), np.float32)
    first = SolveOptions(3, 0, 0, 2, 0.0, 1, 0)
    rest = SolveOptions(1, 12, 4, 2, 0.0, 1, 1)
    for rep in range(2):
        plan.reconstruct_dev(2, d_s, d_a, first)
        plan.reconstruct_dev(2, d_s, d_b, rest, d_m_init=d_a)
        icon.lib().icon_device_synchronize()
        assert rel(d_b.numpy(), full) < 1e-6, rep
        plan.solve(sino, 1, 1, 1, 2, 0.0)   # other work: the next reuse must be refused
        with pytest.raises(icon.IconError):
            plan.reconstruct_dev(2, d_s, d_b, rest, d_m_init=d_a)


def test_warm_start_continues(icon, oracle_mod):
    """A warm start m_init replaces the zero start (Iteration2.cu:33)."""
    tilt, s0 = _small_case(41)
    N = s0.shape[1]
    plan = icon.Plan(N, tilt)
    orc = oracle_mod.Plan(N, N, tilt, 32)
    m_init = orc.solve(s0, 2, 5, 0, 2, 0.0)
    m = plan.solve(s0, 0, 10, 2, 2, 0.0, m_init=m_init)
    mo = orc.solve(s0, 0, 10, 2, 2, 0.0, m_init=m_init)
    assert rel(m, mo) < 1e-4


def test_cross_validation_plan_vs_oracle(icon, oracle_mod):
    """icon_plan_create_cv is the solve without the least-|angle| tilt (the
    other tilts keep their order and get their own weights) reprojected at that
    angle (MainCode_crossV.cu:45-60,113-160): GPU vs the FP32-mirror oracle run
    on the reduced tilt series, on an asymmetric range so that the left-out
    tilt is not the middle one."""
    _print("cross validation plan")
    tilt, sino = _small_case(17)
    tilt, sino = tilt[3:], sino[3:]
    N = sino.shape[1]
    plan = icon.Plan(N, tilt, max_slices=1, cross_validation=True)
    assert plan.info.iomit == len(tilt) - 1
    m, rp, res, al = plan.solve(sino, 4, 20, 4, 2, 0.0, trace=True, reproject=True)
    k = int(np.argmin(np.abs(tilt)))
    orc = oracle_mod.Plan(N, N, np.delete(tilt, k), 32)
    mo, reso, alo = orc.solve(np.delete(sino, k, axis=0), 4, 20, 4, 2, 0.0, trace=True)
    assert rel(m, mo) < 1e-4
    assert np.allclose(res, reso, rtol=1e-4, atol=1e-7)
    assert np.allclose(al, alo, rtol=1e-4)
    rpo = oracle_mod.Plan(N, N, tilt, 32).reproject(mo)
    assert rel(rp, rpo) < 1e-4
    # the full plan's reprojection of the same image is the same line
    full = icon.Plan(N, tilt, max_slices=1)
    assert full.info.iomit == k


def test_cli_end_to_end_with_cross_validation(icon, oracle_mod, tmp_path):
    """icon-gpu on an MRC tilt series: one midYYYYY.mrc per slice equal to the
    oracle's solve, and the cross-validation outputs of ICONGPU.cu:468-592."""
    import subprocess
    from feilab_amd import mrc
    _print("cli end to end")
    rng = np.random.default_rng(23)
    N, S = 64, 5
    tilt = phantom.tilt_series(60, 6)
    b = phantom.blobs_3d(N, S, N, 25, rng)
    sinos = phantom.add_noise(phantom.project_3d(b, N, np.arange(S), tilt), 1.0, rng)
    mrc.write_tilt_series(str(tmp_path / "ts.mrc"), sinos)
    (tmp_path / "ts.tlt").write_text("\n".join("%.2f" % t for t in tilt) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "feilab_amd", "icon-gpu")
    r = subprocess.run([exe, "-i", str(tmp_path / "ts.mrc"), "-t", str(tmp_path / "ts.tlt"), "-o", str(out),
                        "-s", "0,%d" % (S - 1), "-iter", "2,6,2", "-d", "2", "-thr", "0", "-g", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    orc = oracle_mod.Plan(N, N, tilt, 32)
    for s in range(S):
        _, vol = mrc.read(str(out / "reconstruction" / ("mid%05d.mrc" % s)))
        mo = orc.solve(sinos[s], 2, 6, 2, 2, 0.0)
        assert rel(vol[0].T, mo) < 1e-4, s
    cv = out / "crossValidation"
    k = int(np.argmin(np.abs(tilt)))
    _, gt = mrc.read(str(cv / "GroundTruth.mrc"))
    np.testing.assert_array_equal(gt[0], sinos[:, k, :])
    _, full = mrc.read(str(cv / "fullRec_reProjection.mrc"))
    _, cvr = mrc.read(str(cv / "crossV_reprojection.mrc"))
    for name, rec in (("fullRec.frc", full[0]), ("crossV.frc", cvr[0])):
        rows = np.loadtxt(str(cv / name))
        assert rows.shape == (N // 2, 2)
        np.testing.assert_allclose(rows[:, 0], np.arange(N // 2) / N, atol=1e-6)
        expect = icon.icon.frc_smooth(icon.icon.frc(gt[0], rec), 10)
        np.testing.assert_allclose(rows[:, 1], expect, atol=2e-6)
    # the cross-validation reprojection of a slice = the CV plan's, from the oracle
    orc_cv = oracle_mod.Plan(N, N, np.delete(tilt, k), 32)
    mo = orc_cv.solve(np.delete(sinos[1], k, axis=0), 2, 6, 2, 2, 0.0)
    assert rel(cvr[0][1], orc.reproject(mo)) < 1e-4
