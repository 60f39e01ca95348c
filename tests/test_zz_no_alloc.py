"""icon_preprocess_noalloc (no device allocation; modes and tilts staged in the
library's fixed device array) against icon_preprocess on the same stacks:
bitwise, on two calls in a row, the second smaller (fewer images, smaller
images, mode subtraction only), then again the first size after the smaller
call, and the capacity limit refused on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stack(seed, nz, ny, nx):
    rng = np.random.default_rng(seed)
    st = rng.gamma(4.0, 30.0, (nz, ny, nx)) + rng.standard_normal((nz, 1, 1)) * 20
    return st.astype(np.float32)


def _both(icon, st, tilt, thickness):
    from feilab_amd.icon import DeviceArray, preprocess_dev
    nz, ny, nx = st.shape
    d_a = DeviceArray.from_host(st)
    d_b = DeviceArray.from_host(st)
    ma = preprocess_dev(d_a, nz, ny, nx, tilt, thickness)
    mb = preprocess_dev(d_b, nz, ny, nx, tilt, thickness, noalloc=True)
    return d_a.numpy(), ma, d_b.numpy(), mb


def test_noalloc_matches_allocating_entry_point(icon):
    cases = [
        ("first call: 7 images of 64 x 96, normalised", _stack(1, 7, 64, 96), np.linspace(-60, 60, 7), 200),
        ("second call: 3 images of 40 x 50, modes only", _stack(2, 3, 40, 50), None, 0),
        ("third call: 7 images of 64 x 96 again", _stack(3, 7, 64, 96), np.linspace(-45, 50, 7), 120),
    ]
    for name, st, tilt, th in cases:
        print("noalloc preprocess:", name, flush=True)
        out_a, ma, out_b, mb = _both(icon, st, None if tilt is None else tilt.astype(np.float32), th)
        np.testing.assert_array_equal(mb, ma)
        np.testing.assert_array_equal(out_b, out_a)
        assert np.isfinite(out_b).all() and not np.array_equal(out_b, st)


def test_noalloc_refuses_more_images_than_its_scratch(icon):
    from feilab_amd.icon import DeviceArray, IconError, preprocess_dev
    print("noalloc preprocess: capacity check", flush=True)
    d = DeviceArray((1, 4, 4), np.float32)
    with pytest.raises(IconError) as e:
        preprocess_dev(d, 4097, 4, 4, noalloc=True)
    assert e.value.code == 4 and "4096" in str(e.value)
