"""Host-side parts of the product, on the CPU (no device calls):

* the C-ABI library loads and exports every function include/*.h declare;
* the plan geometry uploaded to the device equals the oracle's tables bit for bit;
* MRC reading/writing against the reference's own mrcfile3.c (compiled from
  its source into oracle/_ref by oracle/Makefile, skipped where absent);
* tilt-file reading, the device partition and the command line's error paths.
"""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from feilab_amd import mrc, phantom

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(icon):
    inc = os.path.join(ROOT, "include")
    hdr = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = sorted(set(re.findall(r"\b(icon_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) > 25
    L = icon.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", icon.lib_path()], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r"\bT %s\b" % n, out), n


GEOMS = [(64, 0, phantom.tilt_series(60, 6)), (100, 72, phantom.tilt_series(50, 5)),
         (512, 0, phantom.tilt_series(60, 3)), (333, 200, np.array([-57.3, -41, -2.5, 0.5, 13, 66], np.float32))]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "%dx%d" % (g[0], g[1]))
def test_geometry_equals_oracle_bitwise(icon, oracle_mod, geom):
    nx, thick, tilt = geom
    g = icon.icon.geometry_tables(nx, tilt, thick)
    N1 = thick or nx
    o = oracle_mod.Plan(nx, N1, tilt, 32)
    t = o.tables()
    assert (g["info"].n0, g["info"].n1, g["info"].M, g["info"].iomit) == (o.n0, o.n1, o.M, o.iomit)
    assert g["info"].sigma0 == np.float32(o.sigma[0]) and g["info"].b0 == np.float32(o.b[0])
    np.testing.assert_array_equal(g["nodes"], t["nodes"])
    np.testing.assert_array_equal(g["w"], t["w"])
    np.testing.assert_array_equal(g["phix"], t["phix"].astype(np.float32))
    np.testing.assert_array_equal(g["phiy"], t["phiy"].astype(np.float32))
    np.testing.assert_array_equal(g["ix0"], t["ix"][:, 0])
    np.testing.assert_array_equal(g["iy0"], t["iy"][:, 0])
    # taps are consecutive wrapped cells
    np.testing.assert_array_equal((g["ix0"][:, None] + np.arange(14)) % o.n0, t["ix"])
    np.testing.assert_array_equal(g["ck"], t["ck"].astype(np.float32))
    np.testing.assert_array_equal(g["ck1"], t["ck1"].astype(np.float32))
    np.testing.assert_array_equal(g["phi1"], t["phi1"].astype(np.float32))
    np.testing.assert_array_equal(g["i10"], t["i1"][:, 0])


def test_geometry_rejects_bad_sizes(icon):
    with pytest.raises(icon.IconError):
        icon.icon.geometry_tables(4, phantom.tilt_series(60, 6))
    with pytest.raises(icon.IconError):
        icon.icon.geometry_tables(64, np.array([0.0], np.float32))


# ------------------------------------------------------------------ MRC

def _ref_mrc():
    import oracle
    R = oracle.ref_mrc_lib()
    if R is None:
        pytest.skip("reference mrcfile3.c not built (oracle/_ref); run `make -C oracle ref` with the reference mounted")
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    R.mrc_read_head.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    R.mrc_read_slice.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_char, ctypes.c_void_p]
    R.mrc_update_head.argtypes = [ctypes.c_char_p]
    return R, libc


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
def test_mrc_y_slices_match_reference_reader(icon, tmp_path, mode):
    R, libc = _ref_mrc()
    rng = np.random.default_rng(mode)
    nx, ny, nz = 37, 11, 5
    vol = rng.uniform(0, 200, (nz, ny, nx)) if mode in (0, 6) else rng.uniform(-100, 100, (nz, ny, nx))
    path = str(tmp_path / ("t%d.mrc" % mode))
    mrc.write(path, vol, mode=mode, next_bytes=1024 if mode == 2 else 0)
    mine = icon.icon.mrc_read_y_slices(path, 2, 7)
    hdr = ctypes.create_string_buffer(1024)
    fp = libc.fopen(path.encode(), b"r")
    R.mrc_read_head(fp, hdr)
    for y in range(2, 9):
        buf = np.zeros(nz * nx, np.float32)
        R.mrc_read_slice(fp, hdr, y, b"y", buf.ctypes.data_as(ctypes.c_void_p))
        np.testing.assert_array_equal(mine[y - 2].reshape(-1), buf)
    libc.fclose(fp)


def test_mrc_slice_writer_matches_reference_stats(icon, tmp_path):
    R, _ = _ref_mrc()
    rng = np.random.default_rng(9)
    tpl = str(tmp_path / "tilts.mrc")
    mrc.write(tpl, rng.standard_normal((3, 4, 24)), mode=2, next_bytes=512)
    m = rng.standard_normal((24, 16)).astype(np.float32) * 3 + 1   # [width][thickness]
    out = str(tmp_path / "mid00000.mrc")
    icon.icon.mrc_write_slice(out, tpl, m)
    h, data = mrc.read(out)
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["next"]) == (24, 16, 1, 2, 512)
    assert h["mz"] == 24 and h["zlen"] == 24.0                # mz = mx, zlen = xlen of the input
    np.testing.assert_array_equal(data[0], m.T)              # transposed: x fastest
    ref = str(tmp_path / "ref.mrc")
    shutil.copy(out, ref)
    R.mrc_update_head(ref.encode())
    hr, _ = mrc.read(ref)
    for k in ("amin", "amax", "amean", "rms"):
        assert h[k] == hr[k], k


def test_tilt_file_reader(icon, tmp_path):
    p = tmp_path / "a.tlt"
    p.write_text("-60.0 -57\n  -54.5\n0\n3e1\n")
    np.testing.assert_array_equal(icon.read_tilts(str(p)), np.array([-60, -57, -54.5, 0, 30], np.float32))
    with pytest.raises(icon.IconError) as e:
        icon.read_tilts(str(tmp_path / "missing.tlt"))
    assert "Can not open anglefile" in str(e.value)


# ------------------------------------------------------------- partition

def test_partition_follows_compute_units(icon):
    assert icon.partition_slices([256], 0, 511) == [(0, 511)]
    # floor shares 3 and 6 of 10, the leftover goes to the device with most units
    assert icon.partition_slices([100, 200], 0, 9) == [(0, 2), (3, 9)]
    parts = icon.partition_slices([256, 256, 256, 128], 10, 509)
    assert parts[0][0] == 10 and parts[-1][1] == 509
    assert all(parts[i][1] + 1 == parts[i + 1][0] for i in range(len(parts) - 1))
    # more devices than slices: one slice each
    assert icon.partition_slices([1, 1, 1, 1], 5, 6) == [(5, 5), (6, 6)]


def test_partition_ties_follow_the_reference_loop(icon):
    """Leftover slices go out in the reference's device order (ICONGPU.cu:
    402-414), worked by hand: each pick starts at the last device not yet
    ordered and moves only to a strictly larger count, so 8 equal devices are
    ordered 7, 6, ..., 0 and the two leftovers of 10 slices go to devices 7
    and 6 (device 7, the last, takes the rest)."""
    assert icon.partition_slices([1] * 8, 0, 9) == [(0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (6, 7), (8, 9)]
    # counts 1, 2, 2, 1 and 11 slices: floors 1, 3, 3, 1; the order is 1, 2, 3, 0,
    # so the three leftovers go to devices 1, 2 and 3
    assert icon.partition_slices([1, 2, 2, 1], 0, 10) == [(0, 0), (1, 4), (5, 8), (9, 10)]
    assert icon.partition_slices([256, 256, 256, 128], 0, 499) == [(0, 142), (143, 285), (286, 428), (429, 499)]


# ---------------------------------------------------------------- CLI

def _cli():
    exe = os.path.join(ROOT, "feilab_amd", "icon-gpu")
    if not os.path.exists(exe):
        pytest.fail("icon-gpu not built")
    return exe


def test_cli_help_and_missing_flags():
    r = subprocess.run([_cli(), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "-ICONIteration (-iter)" in r.stdout
    r = subprocess.run([_cli(), "-i", "x.mrc", "-t", "x.tlt"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "parameter error!" in r.stdout


def test_cli_without_device_exits_cleanly(tmp_path):
    """No usable device: the reference prints a notice and exits 0 (ICONGPU.cu:343-357)."""
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([_cli(), "-i", "x.mrc", "-t", "x.tlt", "-o", str(tmp_path), "-s", "0,3", "-iter", "1,1,1"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "No GPU is specified to use" in r.stdout


def test_frc_matches_independent_fft_implementation(icon):
    """icon_frc (host DFT) against the numpy-FFT FRC in feilab_amd.frc, on a
    non-square pair, plus the closed forms frc(a, a) = 1 and frc(a, -a) = -1."""
    from feilab_amd import frc as nfrc
    rng = np.random.default_rng(5)
    a = rng.standard_normal((24, 40)).astype(np.float32)
    b = (a + 0.7 * rng.standard_normal((24, 40))).astype(np.float32)
    c = icon.icon.frc(a, b)
    assert c.shape == (20,)
    np.testing.assert_allclose(c, nfrc.frc(a, b), atol=2e-6)
    np.testing.assert_allclose(icon.icon.frc(a, a), 1.0, atol=1e-6)
    np.testing.assert_allclose(icon.icon.frc(a, -a)[1:], -1.0, atol=1e-6)


def test_frc_smooth_window():
    """frcSmooth: a window of `overlap` entries starting overlap/2 before i,
    shifted inwards at the ends (the last entry is never averaged in)."""
    import feilab_amd
    x = np.arange(8, dtype=np.float32)
    s = feilab_amd.icon.frc_smooth(x, 4)
    # i=0: [0,4) -> 1.5; i=1: [0,4); i=2: [0,4); i=3: [1,5) -> 2.5; i=4: [2,6) -> 3.5;
    # i=5: [3,7) -> 4.5; i=6: [4,8) clipped to [3,7) -> 4.5; i=7: same
    np.testing.assert_allclose(s, [1.5, 1.5, 1.5, 2.5, 3.5, 4.5, 4.5, 4.5])
    np.testing.assert_allclose(feilab_amd.icon.frc_smooth(np.full(30, 0.25, np.float32), 10), 0.25)


def test_preprocess_cli_help_and_missing_flags():
    exe = os.path.join(ROOT, "feilab_amd", "icon-preprocess")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "-thickness, -th" in r.stdout
    r = subprocess.run([exe, "-i", "x.mrc"], capture_output=True, text=True, timeout=30)
    assert "parameter error!" in r.stdout


def test_mask3_host_weights_match_oracle(icon):
    """The library's ICONMask3 mask construction (host C++) against the
    restatement, for the FRC-radius and the given-radius rules."""
    from oracle import mask3_ref
    tilts = np.array([-58.2, -55.1, -51.9, -49.0] + list(np.arange(-45, 46, 3.1)) + [48.7, 52.2, 57.9], np.float32)
    big, mean = mask3_ref.tilt_stats(tilts)
    for N, args in ((128, (0.31, 0.52, 0, -1)), (256, (22 / 128, 0.0, 1, -1)), (128, (0.2, 0.4, 0, 7)),
                    (100, (0.25, 0.45, 0, -1))):
        got = icon.icon.mask3_weights(N, tilts, args[0], args[1], args[2], args[3], mean)
        want = mask3_ref.weights(N, tilts, args[0], args[1], bool(args[2]), args[3], mean)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-7)


def test_mask3_cli_help_and_missing_flags():
    exe = os.path.join(ROOT, "feilab_amd", "icon-mask3")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "-crossVfrc, -cf" in r.stdout
    r = subprocess.run([exe, "-i", "d", "-s", "0,3"], capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "parameter error!" in r.stdout and "-tiltfile" in r.stdout
    # a mask needs '-radius' or both FRC files
    r = subprocess.run([exe, "-i", "d", "-s", "0,3", "-th", "8", "-t", "a.tlt", "-o", "o.mrc"],
                       capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "needs -radius, or both -crossVfrc and -fullRecfrc" in r.stdout
