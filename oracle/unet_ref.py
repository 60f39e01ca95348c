"""CPU oracle of the warm-start U-Net (TEST INFRASTRUCTURE ONLY).

Float64 torch-CPU restatement of the network in feilab_amd/csrc/unet.cpp,
reading the same flat parameter vector.  With bf16=True it rounds where the
device rounds (input, 3x3 weights, every convolution/pool output, and the
output gradient of every 3x3 convolution before its input- and weight-gradient
products) so that the remaining differences are accumulation order and FP32
rounding.  Used only by tests to check the device forward pass and gradients.
"""
import numpy as np
import torch
import torch.nn.functional as F

SPEC = [(8, 32, 1), (32, 32, 32), (32, 64, 32), (64, 64, 64), (64, 128, 64), (128, 128, 128), (128, 256, 128),
        (256, 256, 256), (384, 128, 384), (128, 128, 128), (192, 64, 192), (64, 64, 64), (96, 32, 96), (32, 32, 32)]


def param_count():
    n = sum(9 * co * ci + co for ci, co, _ in SPEC)
    return n + 33


def _bf16(t):
    # round to bfloat16 (nearest even), keep float64; gradient passes straight through
    r = t.to(torch.bfloat16).to(torch.float64)
    return t + (r - t).detach()


class _ConvBF16Grad(torch.autograd.Function):
    """3x3 convolution whose backward rounds the output gradient to BF16
    before the input- and weight-gradient products, as the device's MFMA
    kernels do; the bias gradient sums the unrounded gradient."""

    @staticmethod
    def forward(ctx, x, w, b):
        ctx.save_for_backward(x, w)
        return F.conv2d(x, w, b, padding=1)

    @staticmethod
    def backward(ctx, g):
        x, w = ctx.saved_tensors
        gr = g.to(torch.bfloat16).to(torch.float64)
        gx = torch.nn.grad.conv2d_input(x.shape, w, gr, padding=1)
        gw = torch.nn.grad.conv2d_weight(x, w.shape, gr, padding=1)
        return gx, gw, g.sum((0, 2, 3))


def unpack(params, requires_grad=False):
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=requires_grad)
    layers, off = [], 0
    for ci, co, _ in SPEC:
        w = p[off:off + 9 * co * ci].view(3, 3, co, ci).permute(2, 3, 0, 1)   # [tap][co][ci] -> [co][ci][ky][kx]
        off += 9 * co * ci
        b = p[off:off + co]
        off += co
        layers.append((w, b))
    hw = p[off:off + 32]
    hb = p[off + 32:off + 33]
    return p, layers, hw, hb


def forward(params, x, bf16=True, requires_grad=False):
    """x: [B][H][W] -> out [B][H][W] (float64 torch), plus the parameter tensor."""
    p, L, hw, hb = unpack(params, requires_grad)
    rnd = _bf16 if bf16 else (lambda t: t)
    xt = torch.tensor(np.asarray(x, np.float64))[:, None]
    xt = F.pad(rnd(xt), (0, 0, 0, 0, 0, 7))   # 8 input channels, 7 zero

    def conv(t, i):
        w, b = L[i]
        y = _ConvBF16Grad.apply(t, rnd(w), b) if bf16 else F.conv2d(t, w, b, padding=1)
        return rnd(F.relu(y))

    pool = lambda t: rnd(F.avg_pool2d(t, 2))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    e1 = conv(conv(xt, 0), 1)
    e2 = conv(conv(pool(e1), 2), 3)
    e3 = conv(conv(pool(e2), 4), 5)
    bb = conv(conv(pool(e3), 6), 7)
    d3 = conv(conv(torch.cat([up(bb), e3], 1), 8), 9)
    d2 = conv(conv(torch.cat([up(d3), e2], 1), 10), 11)
    d1 = conv(conv(torch.cat([up(d2), e1], 1), 12), 13)
    out = F.conv2d(d1, hw.view(1, 32, 1, 1), hb)
    return out[:, 0], p


def gradients(params, x, target, bf16=True):
    out, p = forward(params, x, bf16, requires_grad=True)
    t = torch.tensor(np.asarray(target, np.float64))
    loss = ((out - t) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy()
