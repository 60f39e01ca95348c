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

The fallback is per slice: LearnedSolver.solve checks every slice's final
data residual against its target and re-solves the slices that miss it with
plain ICON from m = 0, replacing their image, reprojection and residual.
"""
import math

import numpy as np

from .icon import DeviceArray, Event, Plan, SolveOptions, Stream, copy_d2d, lib
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
    m_start = np.asarray(m_start, np.float32)
    m_final = np.asarray(m_final, np.float32)
    n = m_start.shape[1] * m_start.shape[2]
    s = np.sqrt(np.square(m_start).sum(axis=(1, 2), dtype=np.float64) / n)
    s[s == 0] = 1.0
    inv = (1.0 / s).astype(np.float32)[:, None, None]
    x = m_start * inv
    t = np.subtract(m_final, m_start)
    t *= inv
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


def make_pairs(plan, sino_chunks, total, start, schedule, data_type=2, threshold=0.0, log=None):
    """Normalised training pairs (x, t), float32 [total][N0][N1], from tilt
    series chunks ([s][nz][N0] arrays, s <= plan.max_slices, `total` slices
    in all): each chunk is solved on the device to its start and final
    iterates, normalised and stored, so only one chunk's iterates are held
    beyond the pairs themselves."""
    x = np.empty((total, plan.N0, plan.N1), np.float32)
    t = np.empty((total, plan.N0, plan.N1), np.float32)
    k, nchunk = 0, 0
    for sino in sino_chunks:
        sino = np.ascontiguousarray(sino, np.float32)
        S = sino.shape[0]
        if k + S > total:
            raise ValueError("the chunks hold more than %d slices" % total)
        d_s = DeviceArray.from_host(sino)
        m0, m1 = iterate_pair(plan, d_s, S, start, schedule, data_type, threshold)
        d_s.free()
        x[k:k + S], t[k:k + S] = normalise(m0, m1)
        k += S
        nchunk += 1
        if log and nchunk % 16 == 0:
            log("pairs: %d of %d" % (k, total))
    if k != total:
        raise ValueError("the chunks held %d slices, not %d" % (k, total))
    return x, t


def train_on_pairs(x, t, steps=1500, patch=128, batch=16, lr=1e-3, seed=0, device=0, log=None):
    """Train a U-Net on normalised pairs (make_pairs); returns (params, losses)."""
    net = UNet(patch, patch, max_batch=batch, train=True, device=device, seed=seed + 1)
    tr = Trainer(net, x, t, patch, batch, seed)
    tr.train(steps, lr)
    if log:
        log("trained %d steps on %d pairs: loss %.4g -> %.4g"
            % (steps, x.shape[0], np.mean(tr.losses[:20]), np.mean(tr.losses[-50:])))
    p = net.params()
    net.close()
    return p, tr.losses


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
    """Learned-start reconstruction of up to max_slices slices on a plan, with
    plain ICON (schedule = (pre, icon, post)) as the per-slice fallback.

    overlap=True splits the slices into two halves solved on two streams, each
    with its own plan and network: the second half starts its plain
    iterations when the first half's are done, so they run while the first
    half's U-Net runs (the iterations are VALU- and LDS-bound, the U-Net
    matrix-core-bound), and the first half's continuation overlaps the
    second half's U-Net.  Results are those of the serial run: every slice
    goes through the same kernels."""

    def __init__(self, plan, params, max_slices, start=10, schedule=(20, 160, 20), data_type=2, threshold=0.0,
                 device=0, batch=128, overlap=False):
        self.plan, self.start, self.data_type, self.threshold = plan, start, data_type, threshold
        self.schedule = tuple(int(v) for v in schedule)
        self.overlap = bool(overlap) and max_slices >= 2
        half = (max_slices + 1) // 2
        self.net = UNet(plan.N0, plan.N1, max_batch=min(batch, half if self.overlap else max_slices), device=device)
        self.net.set_params(params)
        self.gain = 1.0
        shape = (max_slices, plan.N0, plan.N1)
        self.d_start = DeviceArray(shape, np.float32)
        self.d_init = DeviceArray(shape, np.float32)
        iters = sum(self.schedule)
        self.d_fb_sino = DeviceArray((max_slices, plan.nz, plan.N0), np.float32)
        self.d_fb_m = DeviceArray(shape, np.float32)
        self.d_fb_rp = DeviceArray((max_slices, plan.N0), np.float32)
        self.d_fb_res = DeviceArray((max_slices, iters + 1), np.float64)
        if self.overlap:
            self.plan2 = Plan(plan.N0, plan.tilt, thickness=plan.thickness, max_slices=half, device=device,
                              cross_validation=plan.cross_validation, window=plan.window)
            if plan.info.col_block != self.plan2.info.col_block:
                self.plan2.set_col_block(plan.info.col_block)
            self.net2 = UNet(plan.N0, plan.N1, max_batch=min(batch, half), device=device)
            self.net2.set_params(params)
            self.streams = (Stream(), Stream())
            self.started = Event()

    def _chain(self, plan, net, S, d_sino, d_start, d_init, d_m, cont, d_resid, d_reproj, stream, use_net,
               after_start=None):
        plan.reconstruct_dev(S, d_sino, d_start, SolveOptions(self.start, 0, 0, self.data_type, self.threshold, 1, 0),
                             stream=stream)
        if after_start is not None:
            after_start()
        if use_net:
            net.warm_start_dev(S, d_start, d_init, stream)
        plan.reconstruct_dev(S, d_sino, d_m, SolveOptions(cont, 0, 0, self.data_type, self.threshold, 1, 1),
                             d_m_init=d_init if use_net else d_start, d_resid=d_resid, d_reproj=d_reproj,
                             stream=stream)

    def set_gain(self, gain):
        """Factor on the network's warm start (icon_unet_set_output_gain), for
        the network(s) of this solver."""
        self.gain = float(gain)
        for net in (self.net, getattr(self, "net2", None)):
            if net is not None:
                net.set_output_gain(self.gain)

    def run(self, S, d_sino, d_m, cont, d_resid=None, d_reproj=None, stream=None, use_net=True):
        """start plain iterations -> U-Net -> `cont` descent iterations into d_m;
        d_resid [S][cont + 1] receives R(m) along the continuation, d_reproj
        [S][N0] the reprojections of the result.  use_net=False continues from
        the start iterate itself (the same method without the network)."""
        if not self.overlap or S < 2:
            self._chain(self.plan, self.net, S, d_sino, self.d_start, self.d_init, d_m, cont, d_resid, d_reproj,
                        stream, use_net)
            return
        p = self.plan
        S0 = (S + 1) // 2
        S1 = S - S0
        sino_b, img_b, res_b, rp_b = p.nz * p.N0 * 4, p.N0 * p.N1 * 4, (cont + 1) * 8, p.N0 * 4
        at = lambda a, off: None if a is None else a.addr(off)
        st0, st1 = self.streams
        if stream is not None:
            ev = Event()
            lib().icon_event_record(ev.h, _stream_handle(stream))
            st0.wait(ev)
            st1.wait(ev)
        self._chain(p, self.net, S0, d_sino, self.d_start, self.d_init, d_m, cont, d_resid, d_reproj, st0, use_net,
                    after_start=lambda: st0.record(self.started))
        st1.wait(self.started)
        self._chain(self.plan2, self.net2, S1, at(d_sino, S0 * sino_b), at(self.d_start, S0 * img_b),
                    at(self.d_init, S0 * img_b), at(d_m, S0 * img_b), cont, at(d_resid, S0 * res_b),
                    at(d_reproj, S0 * rp_b), st1, use_net)
        if stream is None:
            st0.synchronize()
            st1.synchronize()
        else:
            for st in (st0, st1):
                ev = Event()
                st.record(ev)
                lib().icon_stream_wait_event(_stream_handle(stream), ev.h)

    def solve(self, S, d_sino, d_m, cont, target, d_resid, d_reproj=None):
        """The learned solve with its fallback: run(), then every slice whose
        final R exceeds target[s] is re-solved by plain ICON from m = 0 and
        its image, reprojection and final residual (d_resid[s][cont]) are
        replaced.  Synchronises; returns the indices of the re-solved slices."""
        self.run(S, d_sino, d_m, cont, d_resid=d_resid, d_reproj=d_reproj)
        R = d_resid.numpy().reshape(-1)[:S * (cont + 1)].reshape(S, cont + 1)[:, cont]
        bad = np.nonzero(R > np.asarray(target, np.float64)[:S])[0]
        if bad.size:
            self._fallback(bad, d_sino, d_m, cont, d_resid, d_reproj)
            lib().icon_device_synchronize()
        return bad

    def _fallback(self, bad, d_sino, d_m, cont, d_resid, d_reproj):
        p = self.plan
        pre, icon_it, post = self.schedule
        iters = pre + icon_it + post
        row, img, line = p.nz * p.N0 * 4, p.N0 * p.N1 * 4, p.N0 * 4
        for i, s in enumerate(bad):
            copy_d2d(self.d_fb_sino.addr(i * row), d_sino.addr(int(s) * row), row)
        # replayed as a graph: one capture per distinct count of fallback slices
        p.reconstruct_dev(len(bad), self.d_fb_sino, self.d_fb_m,
                          SolveOptions(pre, icon_it, post, self.data_type, self.threshold, 1, 0),
                          d_reproj=self.d_fb_rp if d_reproj is not None else None, d_resid=self.d_fb_res)
        for i, s in enumerate(bad):
            s = int(s)
            copy_d2d(d_m.addr(s * img), self.d_fb_m.addr(i * img), img)
            if d_reproj is not None:
                copy_d2d(d_reproj.addr(s * line), self.d_fb_rp.addr(i * line), line)
            copy_d2d(d_resid.addr((s * (cont + 1) + cont) * 8), self.d_fb_res.addr((i * (iters + 1) + iters) * 8), 8)

    def close(self):
        self.net.close()
        for a in (self.d_start, self.d_init, self.d_fb_sino, self.d_fb_m, self.d_fb_rp, self.d_fb_res):
            a.free()
        if self.overlap:
            self.net2.close()
            self.plan2.close()
            for st in self.streams:
                st.close()
            self.started.close()


def _stream_handle(stream):
    h = getattr(stream, "h", None)
    return h if h is not None else stream


def calibrate_continuation(solver, S, d_sino, target, frc_target=None, sino=None, frc_tol=0.005, max_cont=40,
                           band=None, margin=0.0):
    """Smallest continuation k for which every slice of a calibration batch
    reaches R <= (1 - margin) target[s] (and, given frc_target, the stack's
    reprojection FRC over `band` lies within frc_tol of it); None if no
    k <= max_cont does."""
    p = solver.plan
    d_r = DeviceArray((S, max_cont + 1), np.float64)
    d_rp = DeviceArray((S, p.N0), np.float32)
    d_m = DeviceArray((S, p.N0, p.N1), np.float32)
    try:
        for k in range(1, max_cont + 1):
            solver.run(S, d_sino, d_m, k, d_resid=d_r, d_reproj=d_rp)
            lib().icon_device_synchronize()
            R = d_r.numpy().reshape(-1)[:S * (k + 1)].reshape(S, k + 1)[:, k]
            if not (R <= (1.0 - margin) * np.asarray(target)[:S]).all():
                continue
            if frc_target is not None:
                _, f = reprojection_frc(d_rp.numpy(), sino, p.iomit, band)
                if abs(f - frc_target) > frc_tol:
                    continue
            return k
        return None
    finally:
        for a in (d_r, d_rp, d_m):
            a.free()


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
