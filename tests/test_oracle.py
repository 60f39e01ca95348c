"""Pins the CPU oracle (the CPU restatement of the reference algorithm) with
checks that do not restate its formulas: exact transforms, closed forms,
invariants and special cases that reduce to textbook routines.  Each check is
chosen so a plausible slip (a dropped factor, a wrong sign or index, a
transposed operand) fails it.  CPU only.
"""
import math

import numpy as np
import pytest

from feilab_amd import phantom

from exact_ref import centred_dft_rows, consistent_case, ndft_np, nodes_np


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


# ---------------------------------------------------------------- geometry

def test_grid_is_twice_next_power_of_two(oracle_mod):
    """MainCode.cu:66-71 in float arithmetic gives n = 2 nextpow2(N) for every width."""
    for N in list(range(16, 1100)) + [2047, 2048, 4095, 4096]:
        s, n = oracle_mod.grid(N)
        assert n == 2 * (1 << int(np.ceil(np.log2(N)))), N
        assert abs(s - n / N) < 1e-6 * s


def test_window_shape_closed_form(oracle_mod):
    for sigma in (2.0, 2.56, 3.76):
        b, bf = oracle_mod.window_params(sigma)
        assert b == pytest.approx(2 * sigma * 6 / ((2 * sigma - 1) * np.pi), rel=1e-6)
        assert bf == pytest.approx(1 / np.sqrt(np.pi * b), rel=1e-6)


def test_weights_uniform_spacing_properties(oracle_mod):
    """For uniform angular spacing d every tilt row is the same ramp 2 d |j - c|
    (ends: 2|d|, interior: |2d|), with 180/(4 nz) at the centre."""
    tilt = -phantom.tilt_series(60, 3)          # negated angles, as the reference uses them
    w = oracle_mod.weight(64, tilt)
    j = np.arange(64)
    ramp = 2 * 3.0 * np.abs(j - 32)
    ramp[32] = 180.0 / (4 * len(tilt))
    for i in range(len(tilt)):
        np.testing.assert_allclose(w[i], ramp, rtol=1e-6)


def test_weights_nonuniform_use_neighbour_span(oracle_mod):
    th = np.array([-60, -50, -45, -30, 0, 40], np.float32)
    w = oracle_mod.weight(16, th)
    c = 8
    assert w[0, 0] == pytest.approx(2 * 10 * 8)      # end: twice the gap to the neighbour
    assert w[5, 0] == pytest.approx(2 * 40 * 8)
    assert w[2, 3] == pytest.approx(20 * 5)          # interior: |theta_{i-1} - theta_{i+1}|
    assert w[3, c] == pytest.approx(180.0 / 24)


def test_radial_nodes_on_lines_through_origin(oracle_mod):
    th = np.array([-60, -13.5, 0, 7, 45], np.float32)
    N = 37
    xy = oracle_mod.px2d(N, th).reshape(len(th), N, 2)
    f = -0.5 + np.arange(N) / N
    for i, t in enumerate(th):
        r = np.hypot(xy[i, :, 0], xy[i, :, 1])
        np.testing.assert_allclose(r, np.abs(f), atol=2e-6)
        ang = np.rad2deg(np.arctan2(xy[i, :, 1], xy[i, :, 0]))
        k = np.abs(f) > 0.05
        # direction t (radius > 0) or t - 180 (radius < 0)
        d = (ang[k] - t + 90) % 180 - 90
        np.testing.assert_allclose(d, 0, atol=1e-3)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [16, 256, 1024])
def test_fft_matches_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(oracle_mod.fft(x, -1, 64), np.fft.fft(x)) < 1e-13
    assert rel(oracle_mod.fft(x, +1, 64), np.fft.ifft(x) * n) < 1e-13
    assert rel(oracle_mod.fft(x, -1, 32), np.fft.fft(x)) < 2e-6
    g = rng.standard_normal((8, 32)) + 1j * rng.standard_normal((8, 32))
    assert rel(oracle_mod.fft2(g, -1, 64), np.fft.fft2(g)) < 1e-13


# ---------------------------------------------------------------- NUFFT

@pytest.mark.parametrize("prec,tol", [(64, 1.5e-5), (32, 1.5e-5)])
def test_nufft_matches_exact_ndft(oracle_mod, prec, tol):
    """Gaussian window, m = 6, sigma = 2: error bound 4 exp(-m pi (1 - 1/(2 sigma - 1))) = 1.4e-5."""
    rng = np.random.default_rng(1)
    tilt = phantom.tilt_series(60, 6)
    P = oracle_mod.Plan(32, 32, tilt, prec)
    nodes = P.tables()["nodes"]
    img = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    assert rel(P.forward(img), oracle_mod.ndft_forward(nodes, img)) < tol
    f = rng.standard_normal(P.M) + 1j * rng.standard_normal(P.M)
    assert rel(P.adjoint(f), oracle_mod.ndft_adjoint(nodes, f, 32, 32)) < tol


def test_nufft_adjointness_fp64(oracle_mod):
    rng = np.random.default_rng(2)
    P = oracle_mod.Plan(40, 24, phantom.tilt_series(50, 5), 64)
    x = rng.standard_normal((40, 24)) + 1j * rng.standard_normal((40, 24))
    y = rng.standard_normal(P.M) + 1j * rng.standard_normal(P.M)
    lhs = np.vdot(y, P.forward(x))
    rhs = np.vdot(P.adjoint(y), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_zero_tilt_is_dft_of_column_sums(oracle_mod):
    """At theta = 0 the nodes are (f, 0): the 2D transform reduces to the 1D DFT of
    the sums over the depth axis (Fourier-slice theorem)."""
    rng = np.random.default_rng(3)
    N = 64
    P = oracle_mod.Plan(N, N, np.array([0.0, 30.0], np.float32), 64)
    img = rng.standard_normal((N, N))
    f = P.forward(img.astype(complex))[:N]
    colsum = img.sum(1)
    k = -0.5 + np.arange(N) / N
    X = np.arange(N) - N // 2
    exact = np.exp(-2j * np.pi * np.outer(k, X)) @ colsum
    assert rel(f, exact) < 1.5e-5


def test_projection_transform_is_centred_dft(oracle_mod):
    """Equispaced nodes (k - N/2)/N: the 1D NUFFT of a projection row is the
    centred DFT b_k = sum_X p_X exp(-2 pi i (X - N/2)(k - N/2)/N)."""
    rng = np.random.default_rng(4)
    N = 100
    P = oracle_mod.Plan(N, N, np.array([-10.0, 10.0], np.float32), 64)
    p = rng.standard_normal((2, N)).astype(np.float32)
    b = P.sino_to_b(p)[:N]
    X = np.arange(N) - N // 2
    exact = np.exp(-2j * np.pi * np.outer(X, X) / N) @ p[0].astype(float)
    assert rel(b, exact) < 1.5e-5


# ---------------------------------------------------------------- solver

def _case(seed=5, N=48, step=4, snr=None):
    rng = np.random.default_rng(seed)
    tilt = phantom.tilt_series(60, step)
    b = phantom.blobs_2d(N, 8, rng)
    sino = phantom.project_2d(b, N, tilt)
    if snr:
        sino = phantom.add_noise(sino, snr, rng)
    return tilt, np.asarray(sino, np.float32), b


def test_step_size_is_exact_line_search(oracle_mod):
    """alpha = |rhat|^2 / sum w |A rhat|^2 minimises J(m - alpha rhat) for the complex
    step: recompute the first step from the operators and compare with the trace."""
    tilt, sino, _ = _case()
    P = oracle_mod.Plan(48, 48, tilt, 64)
    _, _, al = P.solve(sino, 1, 0, 0, 2, 0.0, trace=True)
    t = P.tables()
    b = P.sino_to_b(sino)
    w = t["w"].astype(float)
    rhat = -P.adjoint(w * b)                 # gradient at m = 0
    Ar = P.forward(rhat)
    alpha = np.vdot(rhat, rhat).real / np.sum(w * np.abs(Ar) ** 2)
    assert al[0] == pytest.approx(alpha, rel=1e-10)
    # and it is the minimiser: J along the step is stationary there
    J = lambda a: np.sum(w * np.abs(P.forward(-a * rhat) - b) ** 2)
    assert J(alpha) < J(alpha * 1.01) and J(alpha) < J(alpha * 0.99)


@pytest.mark.parametrize("dt,thr", [(2, 0.0), (1, 0.0)])
def test_threshold_bounds_hold(oracle_mod, dt, thr):
    tilt, sino, _ = _case(snr=1.0)
    P = oracle_mod.Plan(48, 48, tilt, 32)
    m = P.solve(sino, 2, 10, 0, dt, thr)
    if dt == 2:
        assert m.min() >= thr
    else:
        assert m.max() <= thr


@pytest.mark.parametrize("prec", [64, 32])
def test_data_residual_matches_exact_transform(oracle_mod, prec):
    """R(m) = sqrt(sum_j w_j |(A m)_j - b_j|^2 / sum_j w_j |b_j|^2) recomputed from
    numpy alone (exact non-uniform DFT of the iterate, centred DFT of the
    projection rows, density weights) after 0, 1, 3 and 10 iterations.  R(0)
    is exactly 1; the NUFFT's window error bounds the rest (1.4e-5 relative
    to |A m|).  A dropped square root, unweighted sums or a wrong
    normalisation are 0.05 .. 0.5 away."""
    N = 32
    rng = np.random.default_rng(5)
    tilt = phantom.tilt_series(60, 6)
    sino = phantom.add_noise(phantom.project_2d(phantom.blobs_2d(N, 6, rng), N, tilt), 1.0, rng)
    P = oracle_mod.Plan(N, N, tilt, prec)
    w = P.tables()["w"].astype(np.float64)
    b = centred_dft_rows(sino)
    xy = nodes_np(N, tilt)
    for k in (0, 1, 3, 10):
        m, res, _ = P.solve(sino, k, 0, 0, 2, 0.0, trace=True)
        assert res[0] == 1.0
        R = np.sqrt(np.sum(w * np.abs(ndft_np(m, xy) - b) ** 2) / np.sum(w * np.abs(b) ** 2))
        assert abs(res[-1] - R) < 2e-6, (k, res[-1], R)
        if k:
            Ru = np.sqrt(np.sum(np.abs(ndft_np(m, xy) - b) ** 2) / np.sum(np.abs(b) ** 2))
            assert abs(Ru - R) > 0.01 and abs(R * R - R) > 0.1


@pytest.mark.parametrize("prec", [64, 32])
def test_data_residual_vanishes_on_consistent_data(oracle_mod, prec):
    """With data that the image m* explains exactly, R(m*) is the window error
    alone (< 1e-5), and INFR from zero drives R towards 0, monotonically,
    well below the noisy case's plateau (about 0.3 at this geometry)."""
    N = 32
    tilt = phantom.tilt_series(60, 6)
    ms, p = consistent_case(N, tilt)
    P = oracle_mod.Plan(N, N, tilt, prec)
    _, res, _ = P.solve(p, 0, 0, 0, 2, 0.0, trace=True, m_init=ms)
    assert res[0] < 1e-5
    _, res, _ = P.solve(p, 200, 0, 0, 2, 0.0, trace=True)
    assert np.all(np.diff(res) <= 1e-12)
    assert res[10] < 0.05 and res[-1] < 0.01


def test_icon_phase_first_clamps_at_pre(oracle_mod):
    """ICON's clamp applies in iterations pre <= it < iters - post
    (Iteration2.cu:129).  With pre = 2, icon = 1 the only clamped step is the
    third, so the result equals the 3-step unclamped solve clamped once; a
    phase starting one step early or late breaks the equality."""
    tilt, sino, _ = _case(snr=1.0)
    P = oracle_mod.Plan(48, 48, tilt, 32)
    unclamped = P.solve(sino, 3, 0, 0, 2, 0.0)
    assert (unclamped < 0).sum() > 50
    np.testing.assert_array_equal(P.solve(sino, 2, 1, 0, 2, 0.0), np.maximum(unclamped, 0.0))
    # dataType 1 clamps from above
    assert (unclamped > 0.01).sum() > 50
    np.testing.assert_array_equal(P.solve(sino, 2, 1, 0, 1, 0.01), np.minimum(unclamped, np.float32(0.01)))


def test_icon_phase_last_clamp_before_post(oracle_mod):
    """The last ICON step is iters - post - 1: pre = 2, icon = 1, post = 1 is the
    clamped 3-step solve followed by one free step (the iteration state is m
    alone, so a warm start from it continues the solve bitwise), which leaves
    values below the threshold; icon = 0 clamps nothing at all."""
    tilt, sino, _ = _case(snr=1.0)
    P = oracle_mod.Plan(48, 48, tilt, 32)
    clamped = P.solve(sino, 2, 1, 0, 2, 0.0)
    full = P.solve(sino, 2, 1, 1, 2, 0.0)
    np.testing.assert_array_equal(full, P.solve(sino, 0, 0, 1, 2, 0.0, m_init=clamped))
    assert (full < 0).sum() > 50
    assert (P.solve(sino, 3, 0, 1, 2, 0.0) < 0).sum() > 50
    np.testing.assert_array_equal(P.solve(sino, 3, 0, 1, 2, 0.0), P.solve(sino, 4, 0, 0, 2, 0.0))


def test_infr_decreases_residual_and_zero_data_stays_zero(oracle_mod):
    tilt, sino, _ = _case()
    P = oracle_mod.Plan(48, 48, tilt, 64)
    _, res, _ = P.solve(sino, 30, 0, 0, 2, 0.0, trace=True)
    assert np.all(np.diff(res) <= 1e-12)
    assert res[-1] < 0.2 * res[0]
    m0 = P.solve(np.zeros_like(sino), 3, 3, 3, 2, 0.0)
    assert np.all(m0 == 0)


def test_fp32_mirror_tracks_fp64(oracle_mod):
    tilt, sino, _ = _case(snr=1.0)
    m32 = oracle_mod.Plan(48, 48, tilt, 32).solve(sino, 3, 20, 3, 2, 0.0)
    m64 = oracle_mod.Plan(48, 48, tilt, 64).solve(sino, 3, 20, 3, 2, 0.0)
    assert rel(m32, m64) < 2e-5


def test_reprojection_is_n_times_projection(oracle_mod):
    """Re(A1D^H A_line m) at the min-|theta| tilt is N0 times that tilt's
    projection (reProject_NFFT.cu:81-98)."""
    N = 64
    tilt, sino, b = _case(N=N, step=6)
    P = oracle_mod.Plan(N, N, tilt, 64)
    truth = phantom.render_2d(b, N, N)
    rp = P.reproject(truth)
    ref = phantom.project_2d(b, N, tilt[[P.iomit]])[0]
    assert np.corrcoef(rp, ref)[0, 1] > 0.999
    assert (rp @ ref) / (ref @ ref) == pytest.approx(N, rel=0.02)


def test_unet_oracle_gradient_matches_finite_differences():
    """The U-Net oracle's backward (float64, no BF16 rounding) against central
    differences of its forward along random directions; with BF16 rounding the
    gradient stays within the rounding-level distance of the exact one."""
    from oracle import unet_ref
    rng = np.random.default_rng(0)
    n = unet_ref.param_count()
    p = rng.standard_normal(n) * 0.05
    x = rng.standard_normal((1, 16, 16))
    t = 0.5 * x
    loss, g = unet_ref.gradients(p, x, t, bf16=False)

    def L(q):
        out, _ = unet_ref.forward(q, x, bf16=False)
        return float(((out.detach().numpy() - t) ** 2).mean())

    for k in range(3):
        d = rng.standard_normal(n)
        d /= np.linalg.norm(d)
        h = 1e-4
        fd = (L(p + h * d) - L(p - h * d)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g @ d)
    _, gb = unet_ref.gradients(p, x, t, bf16=True)
    assert np.linalg.norm(gb - g) / np.linalg.norm(g) < 0.1
    assert float(gb @ g / (np.linalg.norm(gb) * np.linalg.norm(g))) > 0.99


def _central(nx, ny):
    return slice(int(ny * 0.1), ny - int(ny * 0.1)), slice(int(nx * 0.1), nx - int(nx * 0.1))


def test_preprocess_mode_is_lower_edge_of_the_fullest_central_bin(oracle_mod):
    """ICONPreProcess mode: 256 bins over [min, max] of the central region
    (margins excluded); the mode is the lower edge of the fullest bin and is
    subtracted everywhere."""
    rng = np.random.default_rng(3)
    ny, nx = 50, 60
    img = rng.uniform(0.0, 1.0, (ny, nx)).astype(np.float32)
    ys, xs = _central(nx, ny)
    cen = img[ys, xs]
    cen[rng.uniform(size=cen.shape) < 0.2] = np.float32(0.3017)   # a dominant value
    img[ys, xs] = cen
    img[0, :] = 50.0      # margin outliers: ignored
    img[:, -1] = -50.0
    out, modes = oracle_mod.preprocess(img[None])
    mn, mx = float(cen.min()), float(cen.max())
    step = (mx - mn) / 256
    assert 0 <= 0.3017 - modes[0] < step + 1e-12
    assert abs((modes[0] - mn) / step - round((modes[0] - mn) / step)) < 1e-9   # a bin edge
    np.testing.assert_array_equal(out[0], (img.astype(np.float64) - modes[0]).astype(np.float32))


def test_preprocess_ignores_the_maximum_bin_and_constant_images(oracle_mod):
    rng = np.random.default_rng(4)
    ny, nx = 40, 40
    img = rng.uniform(0.0, 0.5, (ny, nx)).astype(np.float32)
    ys, xs = _central(nx, ny)
    cen = img[ys, xs]
    cen[rng.uniform(size=cen.shape) < 0.6] = 1.0      # the maximum, most frequent value
    cen[0, 0] = 0.2                                   # second most frequent cluster
    cen[rng.uniform(size=cen.shape) < 0.1] = 0.2
    img[ys, xs] = cen
    _, modes = oracle_mod.preprocess(img[None])
    step = (1.0 - float(cen.min())) / 256
    assert abs(modes[0] - 0.2) < step          # not the maximum
    out, modes = oracle_mod.preprocess(np.full((2, 20, 30), 7.5, np.float32))
    assert np.all(modes == 7.5) and np.all(out == 0)


def test_preprocess_normalises_variance_to_thickness_over_cosine(oracle_mod):
    rng = np.random.default_rng(5)
    st = (rng.standard_normal((3, 48, 64)) * [[[2.0]], [[0.5]], [[1.0]]] + 3).astype(np.float32)
    tilt = np.array([30.0, 0.0, -45.0], np.float32)
    out, _ = oracle_mod.preprocess(st, tilt, 120)
    want = 0.33 * 120 / np.cos(-tilt / 180.0 * 3.1415926)
    np.testing.assert_allclose(out.reshape(3, -1).astype(np.float64).var(axis=1), want, rtol=1e-5)


def test_mask3_sampling_mask_is_the_tilt_lines():
    """One tilt: the mask is a single line through the centre, perpendicular
    to the rotation by -theta-90 of the central column (theta = 0: the
    central row), with unit values on the line; many tilts cover a wedge and
    stay clipped at 1."""
    from oracle import mask3_ref
    N = 64
    m = mask3_ref.sampling_mask(N, [0.0])
    assert np.allclose(m[N // 2, 1:-1], 1.0, atol=1e-5)
    assert m.sum() == pytest.approx(m[N // 2].sum() + m[N // 2 + 1].sum() + m[N // 2 - 1].sum(), rel=1e-6)
    m90 = mask3_ref.sampling_mask(N, [90.0])
    assert np.allclose(m90[1:-1, N // 2], 1.0, atol=1e-5)
    wedge = mask3_ref.sampling_mask(N, np.arange(-60, 61, 3))
    assert wedge.max() <= 1.0 and wedge[N // 2, N // 2 + 10] == pytest.approx(1.0)
    assert wedge[N // 2 + 20, N // 2] == 0.0            # inside the missing wedge


def test_mask3_weights_closed_forms():
    """realMask = 1 inside the FRC-0.5 radius, the Gaussian edge is 1 at that
    radius and reaches the sampling mask at radius + width; weighting = |k| /
    r_full on sampled frequencies beyond r_full."""
    from oracle import mask3_ref
    N = 128
    tilts = np.arange(-60, 61, 3, dtype=np.float32)
    _, mean = mask3_ref.tilt_stats(tilts)
    r05, r03 = 20 / 64, 40 / 64
    w = mask3_ref.weights(N, tilts, r05, r03, False, -1, mean)
    r_full = math.ceil(2.0 / (float(mean) / 360.0 * 2 * 3.1415926))
    om = mask3_ref.sampling_mask(N, tilts)
    # (ky = 0, kx = r): centred index (N/2, N/2 + r), on a sampled line
    for r in (5, 15, 19):
        assert w[0, r] == pytest.approx(max(1.0, r / r_full) if om[N // 2, N // 2 + r] > 1e-5 else 1.0)
    for r in (45, 60):   # beyond the edge: mask x weighting
        expect = om[N // 2, N // 2 + r] * (r / r_full if (r > r_full and om[N // 2, N // 2 + r] > 1e-5) else 1.0)
        assert w[0, r] == pytest.approx(expect, rel=1e-6)
    # inside the missing wedge (kx = 0, ky = 30): realMask from the edge only
    assert w[30, 0] < 1.0


def test_mask3_radii_from_frc():
    from oracle import mask3_ref
    N = 64
    x = np.arange(N // 2) / (N // 2)
    cv = np.clip(1.2 - 1.5 * x, 0, 1)
    fr = np.ones(N // 2)
    r05, r03, ok = mask3_ref.radii_from_frc(cv, fr, N)
    assert ok and r05 == int(np.nonzero(cv >= 0.5)[0][-1]) and r03 == int(np.nonzero(cv >= 0.3)[0][-1])
    assert not mask3_ref.radii_from_frc(cv[:10], fr[:10], N)[2]        # too short
    only0 = np.zeros(N // 2)
    only0[0] = 0.4
    assert mask3_ref.radii_from_frc(only0, fr, N) == (-1, 0, False)   # FRC-0.3 radius 0
    crossed = np.where(x < 0.5, 0.4, 0.0)
    crossed[:3] = 0.6
    crossed[20:] = 0.0
    assert mask3_ref.radii_from_frc(crossed, fr, N)[2]
    # a curve below 0.3 everywhere gives -1 for both radii, which the reference's check
    # (radius05 > radius03 or radius03 == 0) lets through
    assert mask3_ref.radii_from_frc(np.zeros(N // 2), fr, N) == (-1, -1, True)
