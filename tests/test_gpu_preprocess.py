"""GPU ICONPreProcess (icon_preprocess, icon-preprocess) against the oracle.

Modes are bit-identical (same double arithmetic, integer histogram); the
mode-subtracted images are bit-identical; the normalised images agree to a
few float ulps (the device sums the mean and variance in a different order).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stack(seed, nz=5, ny=64, nx=96):
    rng = np.random.default_rng(seed)
    st = rng.gamma(4.0, 30.0, (nz, ny, nx)) + rng.standard_normal((nz, 1, 1)) * 20
    return st.astype(np.float32)


def test_preprocess_mode_only_matches_oracle(icon, oracle_mod):
    print("preprocess mode", flush=True)
    st = _stack(1)
    out, modes = icon.icon.preprocess(st)
    ref, rmodes = oracle_mod.preprocess(st)
    np.testing.assert_array_equal(modes, rmodes)
    np.testing.assert_array_equal(out, ref)


def test_preprocess_normalised_matches_oracle(icon, oracle_mod):
    print("preprocess normalise", flush=True)
    st = _stack(2)
    tilt = np.linspace(-60, 60, st.shape[0]).astype(np.float32)
    out, modes = icon.icon.preprocess(st, tilt, 200)
    ref, rmodes = oracle_mod.preprocess(st, tilt, 200)
    np.testing.assert_array_equal(modes, rmodes)
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=2e-6 * np.abs(ref).max())


def test_preprocess_cli_end_to_end(icon, oracle_mod, tmp_path):
    """icon-preprocess on a 16-bit MRC stack: float output with the reference's
    header rules, equal to the oracle."""
    from feilab_amd import mrc
    print("preprocess cli", flush=True)
    st = np.round(_stack(3, nz=4, ny=40, nx=50)).astype(np.int16)
    mrc.write(str(tmp_path / "in.mrc"), st, mode=1)
    tilt = np.array([-30.0, -10.0, 10.0, 30.0], np.float32)
    (tmp_path / "in.tlt").write_text("\n".join("%.1f" % t for t in tilt) + "\n")
    exe = os.path.join(ROOT, "feilab_amd", "icon-preprocess")
    r = subprocess.run([exe, "-i", str(tmp_path / "in.mrc"), "-o", str(tmp_path / "out.mrc"), "-t",
                        str(tmp_path / "in.tlt"), "-th", "150"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    h, vol = mrc.read(str(tmp_path / "out.mrc"))
    ref, _ = oracle_mod.preprocess(st.astype(np.float32), tilt, 150)
    np.testing.assert_allclose(vol, ref, rtol=2e-6, atol=2e-6 * np.abs(ref).max())
    assert h["mode"] == 2 and abs(h["amean"] - float(ref.astype(np.float64).mean())) < 1e-3 * max(1.0, ref.std())
