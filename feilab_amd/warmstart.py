"""Learned warm start of the ICON solve.

A plain ICON solve (MainCode.cu:113-135, Iteration2.cu:33-146) starts from
m = 0 and runs pre + icon + post steepest-descent iterations.  The learned
start runs the first `start` iterations, maps that iterate through the U-Net
(icon_unet_warm_start: m + s U(m / s), s the slice RMS) and continues with
unthresholded descent iterations from the prediction until every slice's data
residual R(m) is at or below a target (the residual plain ICON ends with) and
the reprojection FRC at the min-|theta| tilt is within a tolerance of plain
ICON's.  Everything runs on the device through the C ABI; plain ICON
(m_init = NULL) stays the fallback for any slice that misses the target.

Training pairs are (m_start, m_final) from the device solve itself: the
continuation of an iterate under the rest of the schedule is the uninterrupted
solve, because the iteration state is m alone.
"""
import math

import numpy as np

from .icon import DeviceArray, SolveOptions, lib
from .unet import UNet


def iterate_pair(plan, d_sino, S, start, schedule, data_type=2, threshold=0.0):
    """(m_start, m_final) host arrays [S][N0][N1] of the plain schedule
    (pre, icon, post): m_start after `start` (<= pre) iterations, m_final after
    all of them."""
    pre, icon_it, post = schedule
    if not 0 < start <= pre:
        raise ValueError("start must lie in 1..pre")
    d_a = DeviceArray((S, plan.N0, plan.N1), np.float32)
    d_b = DeviceArray((S, plan.N0, plan.N1), np.float32)
    plan.reconstruct_dev(S, d_sino, d_a, SolveOptions(start, 0, 0, data_type, threshold, 1, 0))
    plan.reconstruct_dev(S, d_sino, d_b, SolveOptions(pre - start, icon_it, post, data_type, threshold, 1, 1),
                         d_m_init=d_a)
    lib().icon_device_synchronize()
    return d_a.numpy(), d_b.numpy()


def normalise(m_start, m_final):
    """Network input m_start / s and target (m_final - m_start) / s, s the RMS
    of each m_start slice (1 for a zero slice): the scaling of the warm start."""
    s = np.sqrt((m_start.astype(np.float64) ** 2).mean(axis=(1, 2)))
    s[s == 0] = 1.0
    x = (m_start / s[:, None, None]).astype(np.float32)
    t = ((m_final.astype(np.float64) - m_start) / s[:, None, None]).astype(np.float32)
    return x, t


class Trainer:
    """Adam on random patches (with the flips the symmetric tilt range allows)
    of (input, target) slices, on the device's own forward/backward kernels."""

    def __init__(self, net, x, t, patch=128, batch=16, seed=0):
        if net.max_batch < batch or net.H != patch or net.W != patch:
            raise ValueError("the training net must be patch x patch with max_batch >= batch")
        self.net, self.x, self.t = net, x, t
        self.patch, self.batch = patch, batch
        self.rng = np.random.default_rng(seed)
        self.d_x = DeviceArray((batch, patch, patch), np.float32)
        self.d_t = DeviceArray((batch, patch, patch), np.float32)
        self.losses = []

    def sample(self):
        S, H, W = self.x.shape
        P, B, r = self.patch, self.batch, self.rng
        xs = np.empty((B, P, P), np.float32)
        ts = np.empty((B, P, P), np.float32)
        for k in range(B):
            i, y0, x0 = r.integers(S), r.integers(H - P + 1), r.integers(W - P + 1)
            a = self.x[i, y0:y0 + P, x0:x0 + P]
            b = self.t[i, y0:y0 + P, x0:x0 + P]
            if r.integers(2):      # x -> -x with the tilt range mirrored
                a, b = a[::-1], b[::-1]
            if r.integers(2):      # z -> -z
                a, b = a[:, ::-1], b[:, ::-1]
            xs[k], ts[k] = a, b
        return xs, ts

    def train(self, steps, lr=1e-3, lr_min=5e-5, warmup=50):
        for it in range(steps):
            xs, ts = self.sample()
            self.d_x.upload(xs)
            self.d_t.upload(ts)
            if it < warmup:
                a = lr * (it + 1) / warmup
            else:
                a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * (it - warmup) / max(1, steps - warmup)))
            self.losses.append(self.net.train_step_dev(self.batch, self.d_x, self.d_t, a))
        return self.losses


def train_warm_start(plan, sino, start, schedule, steps=1500, patch=128, batch=16, lr=1e-3, seed=0,
                     data_type=2, threshold=0.0, device=0, log=None):
    """Train a U-Net on (start-iterate, final-iterate) pairs of the sinograms
    [S][nz][N0]; returns (params, losses)."""
    sino = np.ascontiguousarray(sino, np.float32)
    S = sino.shape[0]
    d_s = DeviceArray.from_host(sino)
    m0, m1 = iterate_pair(plan, d_s, S, start, schedule, data_type, threshold)
    x, t = normalise(m0, m1)
    net = UNet(patch, patch, max_batch=batch, train=True, device=device, seed=seed + 1)
    tr = Trainer(net, x, t, patch, batch, seed)
    tr.train(steps, lr)
    if log:
        log("trained %d steps: loss %.4g -> %.4g" % (steps, np.mean(tr.losses[:20]), np.mean(tr.losses[-50:])))
    p = net.params()
    net.close()
    return p, tr.losses


class LearnedSolver:
    """Learned-start reconstruction of up to max_slices slices on a plan."""

    def __init__(self, plan, params, max_slices, start=4, data_type=2, threshold=0.0, device=0, batch=128):
        self.plan, self.start, self.data_type, self.threshold = plan, start, data_type, threshold
        self.net = UNet(plan.N0, plan.N1, max_batch=min(batch, max_slices), device=device)
        self.net.set_params(params)
        shape = (max_slices, plan.N0, plan.N1)
        self.d_start = DeviceArray(shape, np.float32)
        self.d_init = DeviceArray(shape, np.float32)

    def run(self, S, d_sino, d_m, cont, d_resid=None, d_reproj=None, stream=None, use_net=True):
        """start plain iterations -> U-Net -> `cont` descent iterations into d_m;
        d_resid [S][cont + 1] receives R(m) along the continuation, d_reproj
        [S][N0] the reprojections of the result.  use_net=False continues from
        the start iterate itself (the same method without the network)."""
        p = self.plan
        p.reconstruct_dev(S, d_sino, self.d_start, SolveOptions(self.start, 0, 0, self.data_type, self.threshold, 1, 0),
                          stream=stream)
        if use_net:
            self.net.warm_start_dev(S, self.d_start, self.d_init, stream)
        p.reconstruct_dev(S, d_sino, d_m, SolveOptions(cont, 0, 0, self.data_type, self.threshold, 1, 1),
                          d_m_init=self.d_init if use_net else self.d_start, d_resid=d_resid, d_reproj=d_reproj,
                          stream=stream)

    def close(self):
        self.net.close()


def reprojection_frc(reproj, sino, iomit, band):
    """ICON's fullRec FRC (calFRC_nonsqure over the slice stack) of the
    reprojections [S][N0] against the measured projection rows at the
    min-|theta| tilt, averaged over rings band[0] .. band[1]-1."""
    from .frc import frc
    c = frc(np.asarray(reproj, np.float64), np.asarray(sino)[:, iomit, :])
    return c, float(c[band[0]:band[1]].mean())


def iterations_to_target(resid, target):
    """First continuation count k with R_k <= target for every slice
    (resid [S][K + 1], target [S]); None if the trace never gets there."""
    ok = (resid <= np.asarray(target)[:, None]).all(axis=0)
    idx = np.nonzero(ok)[0]
    return int(idx[0]) if idx.size else None
