"""GPU ICONMask3 (icon_mask3_filter, icon-mask3) against the oracle.

The oracle filters in FP64 (numpy rfft2 / irfft2); the device uses FP32
FFTs (radix-16 for power-of-two widths, Bluestein otherwise), so slices agree
to a few 1e-6 of their largest value.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _close(got, want, tol=2e-5):
    scale = float(np.abs(want).max())
    err = float(np.abs(got - want).max())
    assert err <= tol * scale, (err, scale)


@pytest.mark.parametrize("N", [128, 256, 96, 200, 64])
def test_mask3_filter_matches_oracle(icon, N):
    """A random half-spectrum multiplier exercises the mirroring of every
    column, including kx = 0 and kx = N/2."""
    from oracle import mask3_ref
    print("mask3 filter N=%d" % N, flush=True)
    rng = np.random.default_rng(N)
    sl = rng.standard_normal((3, N, N)).astype(np.float32)
    half = rng.uniform(0, 2, (N, N // 2 + 1)).astype(np.float32)
    _close(icon.icon.mask3_filter(sl, half), mask3_ref.filter_slices(sl, half))


def test_mask3_filter_with_tilt_mask(icon):
    from oracle import mask3_ref
    print("mask3 filter tilt mask", flush=True)
    N = 160
    tilts = np.arange(-60, 61, 3.0).astype(np.float32)
    _, mean = mask3_ref.tilt_stats(tilts)
    half = mask3_ref.weights(N, tilts, 0.3, 0.5, False, -1, mean)
    rng = np.random.default_rng(5)
    sl = rng.standard_normal((2, N, N)).astype(np.float32)
    _close(icon.icon.mask3_filter(sl, half), mask3_ref.filter_slices(sl, half))


def _icon_run(tmp_path, N, S, tilt, rng):
    from feilab_amd import mrc, phantom
    b = phantom.blobs_3d(N, S, N, 25, rng)
    sinos = phantom.add_noise(phantom.project_3d(b, N, np.arange(S), tilt), 1.0, rng)
    mrc.write_tilt_series(str(tmp_path / "ts.mrc"), sinos)
    (tmp_path / "ts.tlt").write_text("\n".join("%.2f" % t for t in tilt) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    exe = os.path.join(ROOT, "feilab_amd", "icon-gpu")
    r = subprocess.run([exe, "-i", str(tmp_path / "ts.mrc"), "-t", str(tmp_path / "ts.tlt"), "-o", str(out),
                        "-s", "0,%d" % (S - 1), "-iter", "2,6,2", "-d", "2", "-thr", "0", "-g", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def _mids(out, S):
    from feilab_amd import mrc
    return np.stack([mrc.read(str(out / "reconstruction" / ("mid%05d.mrc" % s)))[1][0] for s in range(S)])


def test_mask3_cli_end_to_end(icon, tmp_path):
    """icon-gpu (with its cross-validation pass) then icon-mask3 with the FRC
    files it wrote, with a given radius, and with no mask: each output volume
    equal to the oracle's filtering and cropping of the midYYYYY.mrc slices."""
    from feilab_amd import mrc, phantom
    from oracle import mask3_ref
    print("mask3 cli", flush=True)
    rng = np.random.default_rng(31)
    N, S, th, zs = 96, 4, 40, 3
    tilt = phantom.tilt_series(60, 6)
    out = _icon_run(tmp_path, N, S, tilt, rng)
    mids = _mids(out, S)
    thita = np.loadtxt(str(tmp_path / "ts.tlt")).astype(np.float32)
    _, mean = mask3_ref.tilt_stats(thita)
    cvf, frf = out / "crossValidation" / "crossV.frc", out / "crossValidation" / "fullRec.frc"
    exe = os.path.join(ROOT, "feilab_amd", "icon-mask3")
    base = [exe, "-i", str(out / "reconstruction"), "-s", "0,%d" % (S - 1), "-th", str(th),
            "-t", str(tmp_path / "ts.tlt"), "-z", str(zs)]
    y0 = N // 2 - th // 2 + zs

    # FRC-derived radius
    cv, fr = np.loadtxt(str(cvf))[:, 1], np.loadtxt(str(frf))[:, 1]
    r05, r03, ok = mask3_ref.radii_from_frc(cv, fr, N)
    if ok:
        half = mask3_ref.weights(N, thita, np.float32(r05) / np.float32(len(fr)),
                                 np.float32(r03) / np.float32(len(fr)), False, -1, mean)
    else:
        big, _ = mask3_ref.tilt_stats(thita)
        half = mask3_ref.weights(N, thita, np.float32(big) / np.float32(N / 2), 0.0, True, -1, mean)
    r = subprocess.run(base + ["-o", str(tmp_path / "frc.mrc"), "-cf", str(cvf), "-ff", str(frf)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    h, vol = mrc.read(str(tmp_path / "frc.mrc"))
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (N, th, S, 2)
    _close(vol, mask3_ref.filter_slices(mids, half)[:, y0:y0 + th, :])

    # given radius
    r = subprocess.run(base + ["-o", str(tmp_path / "rad.mrc"), "-r", "20"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    half = mask3_ref.weights(N, thita, np.float32(20) / np.float32(N / 2), 0.0, True, -1, mean)
    _close(mrc.read(str(tmp_path / "rad.mrc"))[1], mask3_ref.filter_slices(mids, half)[:, y0:y0 + th, :])

    # no mask: crop and stack only; exit status 1 after simpleCatch, as the reference
    r = subprocess.run(base + ["-o", str(tmp_path / "nm.mrc"), "-nm", "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "without the Fourier mask" in r.stdout, r.stdout + r.stderr
    np.testing.assert_array_equal(mrc.read(str(tmp_path / "nm.mrc"))[1], mids[:, y0:y0 + th, :])
