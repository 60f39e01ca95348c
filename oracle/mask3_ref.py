"""CPU restatement of ICONMask3 (TEST INFRASTRUCTURE ONLY).

Follows sourceCode/ICONMask3/ICONCrossVMaskIn2D_weightOnlyOnSampled.cu:
  sampling_mask  main, lines 958-1000 (rotated central line, bilinear splat,
                 float32 accumulation in the reference's order, clip to 1)
  weights        genMask, lines 369-455 (realMask * weighting)
  radii          main, lines 866-940 (FRC-0.5 / FRC-0.3 radii, validation)
  filter         maskAndWeightSliceCU, lines 479-598 (R2C, multiply, C2R / N^2)
                 as numpy rfft2 / irfft2 (FP64)
Used only by tests to check the product's host and device code.
"""
import math

import numpy as np

PI_180 = 0.01745329252


def sampling_mask(N, thita):
    nx = ny = N
    om = np.zeros((ny, nx), np.float32)
    j = np.arange(nx)
    i = np.full(nx, nx // 2)
    for t in thita:
        # (-theta - 90) in float, times PI_180 in double (numpy scalars would stay float32)
        a = float(-np.float32(t) - np.float32(90)) * PI_180
        c = np.float32(math.cos(a))
        s = np.float32(math.sin(a))
        fi = (i - nx // 2).astype(np.float32)
        fj = (j - ny // 2).astype(np.float32)
        xf = (fi * c - fj * s) + np.float32(nx // 2)
        yf = (fi * s + fj * c) + np.float32(ny // 2)
        xi = xf.astype(np.int64)            # C truncation toward zero
        yi = yf.astype(np.int64)
        dx = xf - xi.astype(np.float32)
        dy = yf - yi.astype(np.float32)
        one = np.float32(1)
        vals = np.stack([(one - dx) * (one - dy), (one - dx) * dy, dx * (one - dy), dx * dy], 1)   # per pixel, in order
        ys = np.stack([yi, yi + 1, yi, yi + 1], 1)
        xs = np.stack([xi, xi, xi + 1, xi + 1], 1)
        ok = (xs >= 0) & (xs < nx) & (ys >= 0) & (ys < ny)
        rot = np.zeros((ny, nx), np.float32)
        np.add.at(rot, (ys[ok], xs[ok]), vals[ok].astype(np.float32))
        om += rot
    return np.minimum(om, np.float32(1))


def weights(N, thita, r05_rate, r03_rate, radius_given, gauss_width, mean_step):
    """realMask * weighting on the half spectrum [N][N/2+1]."""
    om = sampling_mask(N, thita)
    r_full = math.ceil(2.0 / (float(mean_step) / 360.0 * 2 * 3.1415926))
    radius05 = int(float(N) / 2.0 * float(np.float32(r05_rate)))
    radius03 = int(float(N) / 2.0 * float(np.float32(r03_rate)))
    if not radius_given:
        glm = gauss_width if gauss_width >= 0 else radius03 - radius05
    else:
        glm = gauss_width if gauss_width >= 0 else int(10.0 * (N / 150.0))
    if glm < 10.0 * (N / 150.0):
        glm = int(10.0 * (N / 150.0))
    radius = radius05
    grad = radius + glm
    gl = grad - radius
    c = gl / 3.0
    e_end = math.exp(-(gl * gl) / (2 * c * c))
    H = N // 2 + 1
    j, i = np.meshgrid(np.arange(N), np.arange(H), indexing="ij")
    i1 = (i + N // 2) % N
    j1 = (j + N // 2) % N
    rr = np.floor(np.sqrt(((i1 - N // 2) ** 2 + (j1 - N // 2) ** 2).astype(np.float64))).astype(np.int64)
    o = om[j1, i1]
    w2 = (np.exp(-((rr - radius).astype(np.float64) ** 2) / (2 * c * c)) - e_end) / (1 - e_end)
    real = np.where(rr >= radius, np.where(rr <= grad, np.maximum(o.astype(np.float64), w2), o), 1.0).astype(np.float32)
    weighting = np.where((rr > r_full) & (np.abs(o) > 0.00001), rr.astype(np.float32) / np.float32(r_full),
                         np.float32(1)).astype(np.float32)
    return (real * weighting).astype(np.float32)


def radii_from_frc(crossv, fullrec, N):
    """(radius05, radius03, valid) from FRC value columns."""
    cv = np.array(crossv, np.float32)
    fr = np.array(fullrec, np.float32)
    if len(cv) < N // 2 - 1 or len(fr) < N // 2 - 1:
        return None, None, False
    ratio = cv.copy()
    nz = fr != 0
    ratio[: len(fr)][nz] = cv[: len(fr)][nz] / fr[nz]

    def last(th):
        idx = np.nonzero(ratio[: len(fr)] >= th)[0]
        return int(idx[-1]) if idx.size else -1
    r05, r03 = last(0.5), last(0.3)
    if r05 > r03 or r03 == 0:
        return r05, r03, False
    return r05, r03, True


def tilt_stats(thita):
    big = np.float32(-1)
    small = np.float32(1)
    for t in thita:
        big = max(big, np.float32(t))
        small = min(small, np.float32(t))
    big_radius = math.ceil(2 * 180.0 / 3.1415926 / float(np.float32(big - small)))
    mean = np.float32(0)
    for a, b in zip(thita[:-1], thita[1:]):
        mean = np.float32(mean + np.float32(abs(np.float32(b) - np.float32(a))))
    return big_radius, np.float32(mean / np.float32(len(thita) - 1))


def filter_slices(slices, half_mask):
    """R2C -> multiply -> C2R / N^2 of [S][N][N] slices (FP64)."""
    F = np.fft.rfft2(np.asarray(slices, np.float64))
    return np.fft.irfft2(F * half_mask.astype(np.float64), s=slices.shape[-2:])
