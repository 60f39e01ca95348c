"""GPU checks of the warm-start U-Net against its CPU oracle (oracle/unet_ref.py).

The device runs BF16 MFMA convolutions with FP32 accumulation; the oracle is
float64 with BF16 rounding at the same points (input, weights, every
activation).  Tolerances:

* forward: relative L2 <= 1e-2 (accumulation order flips an occasional BF16
  rounding; a wrong operand layout or tap gives O(1) errors);
* gradients: cosine >= 0.999 overall, and per parameter block a relative L2
  distance to the BF16 oracle of at most twice the oracle's own BF16-vs-FP64
  distance for that block (1e-2 at least): two BF16 computations that round
  in a different order differ by about the rounding effect itself, while a
  wrong layout, tap or mask gives O(1).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _blocks():
    from oracle import unet_ref
    out, off = [], 0
    for i, (ci, co, _) in enumerate(unet_ref.SPEC):
        out.append(("w%d" % i, off, off + 9 * co * ci))
        off += 9 * co * ci
        out.append(("b%d" % i, off, off + co))
        off += co
    out.append(("head", off, off + 33))
    return out


def _net_and_params(H, W, B, seed):
    from feilab_amd.unet import UNet
    net = UNet(H, W, max_batch=B, train=True, seed=seed)
    p = net.params()
    rng = np.random.default_rng(seed)
    # non-zero biases so that the bias paths are exercised
    for name, a, b in _blocks():
        if name.startswith("b"):
            p[a:b] = 0.05 * rng.standard_normal(b - a)
    p[-1] = 0.1
    net.set_params(p)
    return net, net.params()


@pytest.mark.parametrize("shape", [(1, 16, 16), (2, 32, 48), (3, 64, 40)])
def test_unet_forward_vs_oracle(icon, shape):
    from oracle import unet_ref
    B, H, W = shape
    _print("unet forward", shape)
    net, p = _net_and_params(H, W, B + 1, seed=B * 7 + H)
    assert net.nparam == unet_ref.param_count()
    rng = np.random.default_rng(H * W)
    x = rng.standard_normal(shape).astype(np.float32)
    y = net.forward(x)
    ref, _ = unet_ref.forward(p, x, bf16=True)
    ref = ref.detach().numpy()
    assert np.isfinite(y).all()
    assert rel(y, ref) < 1e-2, rel(y, ref)
    net.close()


@pytest.mark.parametrize("shape", [(1, 16, 16), (2, 32, 32)])
def test_unet_gradients_vs_oracle(icon, shape):
    from oracle import unet_ref
    B, H, W = shape
    _print("unet gradients", shape)
    net, p = _net_and_params(H, W, B, seed=11 + B)
    rng = np.random.default_rng(5 + B)
    x = rng.standard_normal(shape).astype(np.float32)
    t = (0.5 * x + 0.1 * rng.standard_normal(shape)).astype(np.float32)
    loss, g = net.gradients(x, t)
    rloss, rg = unet_ref.gradients(p, x, t, bf16=True)
    _, rg64 = unet_ref.gradients(p, x, t, bf16=False)
    assert abs(loss - rloss) <= 1e-2 * abs(rloss)
    gd_all = g.astype(np.float64)
    assert float(gd_all @ rg / (np.linalg.norm(gd_all) * np.linalg.norm(rg))) > 0.999
    worst = []
    for name, a, b in _blocks():
        gd, gr, g64 = g[a:b].astype(np.float64), rg[a:b], rg64[a:b]
        if np.linalg.norm(gr) == 0:
            assert np.linalg.norm(gd) == 0, name
            continue
        # BF16 rounding alone moves this block by `noise` (oracle BF16 vs FP64);
        # two BF16 computations with different rounding order differ by about as much
        noise = rel(gr, g64)
        worst.append((rel(gd, gr) / max(noise, 5e-3), rel(gd, gr), noise, name))
        assert rel(gd, gr) <= max(2.0 * noise, 1e-2), (name, rel(gd, gr), noise)
    _print("worst block (ratio, rel, bf16 noise, name)", max(worst))
    net.close()


def test_unet_train_step_lowers_loss(icon):
    from feilab_amd.icon import DeviceArray
    from feilab_amd.unet import UNet
    _print("unet train steps")
    net = UNet(32, 32, max_batch=4, train=True, seed=1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 32, 32)).astype(np.float32)
    t = np.roll(x, 1, axis=2) * 0.5
    dx, dt = DeviceArray.from_host(x), DeviceArray.from_host(t.astype(np.float32))
    losses = [net.train_step_dev(4, dx, dt, 1e-3) for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0], losses[::10]
    # the BF16 compute copy follows the master weights: forward agrees with the oracle afterwards
    from oracle import unet_ref
    y = net.forward(x)
    ref = unet_ref.forward(net.params(), x, bf16=True)[0].detach().numpy()
    assert rel(y, ref) < 1e-2
    net.close()


def test_unet_rejects_bad_sizes(icon):
    from feilab_amd.unet import UNet, UNetError
    with pytest.raises(UNetError):
        UNet(36, 32)
    net = UNet(16, 16, max_batch=1)
    with pytest.raises(UNetError):
        net.forward(np.zeros((2, 16, 16), np.float32))
    with pytest.raises(UNetError):
        net.gradients(np.zeros((1, 16, 16), np.float32), np.zeros((1, 16, 16), np.float32))
    net.close()


def test_unet_warm_start_is_scaled_residual(icon):
    """warm_start(m) = m + s U(m / s) with s the RMS of each slice, in batches of max_batch."""
    from feilab_amd.unet import UNet
    _print("unet warm start")
    net = UNet(32, 40, max_batch=2, seed=4)
    rng = np.random.default_rng(9)
    m = (rng.standard_normal((5, 32, 40)) * np.array([1e-3, 1.0, 30.0, 0.2, 5.0])[:, None, None]).astype(np.float32)
    m[3] = 0.0
    out = net.warm_start(m)
    s = np.sqrt((m.astype(np.float64) ** 2).mean(axis=(1, 2)))
    s[s == 0] = 1.0
    y = np.concatenate([net.forward((m[i:i + 2] / s[i:i + 2, None, None]).astype(np.float32)) for i in (0, 2)] +
                       [net.forward((m[4:5] / s[4]).astype(np.float32))])
    ref = m + s[:, None, None] * y
    for i in range(5):
        assert rel(out[i], ref[i]) < 2e-3, i   # host and device round m / s to BF16 separately
    # the output gain multiplies the result (one FP32 product per pixel), and
    # gain 1 is the plain result again
    net.set_output_gain(0.85)
    assert np.array_equal(net.warm_start(m), (np.float32(0.85) * out).astype(np.float32))
    net.set_output_gain(1.0)
    assert np.array_equal(net.warm_start(m), out)
    from feilab_amd.unet import UNetError
    with pytest.raises(UNetError):
        net.set_output_gain(0.0)
    net.close()


def test_unet_inference_net_equals_training_net(icon):
    """The inference net reads its decoders' upsampled inputs straight from
    the half-size tensors; the training net (checked against the oracle above)
    materialises them.  Same parameters, same input: identical output."""
    from feilab_amd.unet import UNet
    _print("unet inference vs training path")
    tr, p = _net_and_params(64, 48, 3, seed=21)
    inf = UNet(64, 48, max_batch=3)
    inf.set_params(p)
    x = np.random.default_rng(5).standard_normal((3, 64, 48)).astype(np.float32)
    np.testing.assert_array_equal(inf.forward(x), tr.forward(x))
    inf.close()
    tr.close()
