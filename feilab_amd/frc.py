"""Fourier ring correlation, as ICON-GPU computes it (calFRC.cu:8-174).

Rings r = floor(sqrt(dx^2 + (dy * nx/ny)^2)) in the centred spectrum of an
nx-wide, ny-high image (nx >= ny), r < nx/2; frc[0] = 1, frc[r] =
Re sum F1 conj(F2) / sqrt(sum |F1|^2 sum |F2|^2), 0 for an empty ring.  The
transforms run on the host (numpy, FP64): the FRC is an evaluation of finished
reconstructions, not part of the iteration.
"""
import numpy as np


def _rings(nx, ny):
    # the reference's float arithmetic: (float)(di^2) + (float)(dj^2) * damp2
    # with damp2 = ((float)nx / (float)ny)^2 in float, sqrt in double
    i = (np.arange(nx) + nx // 2) % nx - nx // 2
    j = (np.arange(ny) + ny // 2) % ny - ny // 2
    d = np.float32(np.float32(nx) / np.float32(ny))
    damp2 = np.float32(d * d)
    q = (i[None, :] ** 2).astype(np.float32) + (j[:, None] ** 2).astype(np.float32) * damp2
    r = np.floor(np.sqrt(q.astype(np.float64))).astype(np.int64)
    return r


def frc(image1, image2):
    """FRC curve (length nx/2) of two images [ny][nx] (x fastest), or of two
    stacks [..., ny, nx] with the ring sums pooled over the stack."""
    a = np.asarray(image1, np.float64)
    b = np.asarray(image2, np.float64)
    ny, nx = a.shape[-2:]
    if nx < ny:
        raise ValueError("frc: width must be at least the height (calFRC_nonsqure)")
    A = np.fft.fft2(a)
    B = np.fft.fft2(b)
    r = _rings(nx, ny)
    keep = r < nx // 2
    idx = r[keep]
    up = (A * np.conj(B)).real.reshape(-1, ny, nx)[:, keep].sum(0)
    l1 = (np.abs(A) ** 2).reshape(-1, ny, nx)[:, keep].sum(0)
    l2 = (np.abs(B) ** 2).reshape(-1, ny, nx)[:, keep].sum(0)
    U = np.bincount(idx, up, nx // 2)
    L1 = np.bincount(idx, l1, nx // 2)
    L2 = np.bincount(idx, l2, nx // 2)
    den = np.sqrt(L1) * np.sqrt(L2)
    out = np.where(den != 0, U / np.where(den != 0, den, 1), 0.0)
    out[0] = 1.0
    return out


def frc_slices(images1, images2):
    """FRC curve of every pair of slices of two stacks [S][ny][nx]: [S][nx/2]
    (each slice on its own, not pooled as frc() pools a stack)."""
    a = np.asarray(images1, np.float32)
    b = np.asarray(images2, np.float32)
    S, ny, nx = a.shape
    if nx < ny:
        raise ValueError("frc: width must be at least the height (calFRC_nonsqure)")
    r = _rings(nx, ny)
    keep = r < nx // 2
    idx = r[keep]
    nr = nx // 2
    out = np.zeros((S, nr))
    try:
        from scipy import fft as sfft
        fft2 = lambda x: sfft.fft2(x, workers=-1)
    except ImportError:
        fft2 = np.fft.fft2
    for s0 in range(0, S, 16):
        A = fft2(a[s0:s0 + 16].astype(np.float64))
        B = fft2(b[s0:s0 + 16].astype(np.float64))
        k = A.shape[0]
        up = (A * np.conj(B)).real[:, keep]
        l1 = (np.abs(A) ** 2)[:, keep]
        l2 = (np.abs(B) ** 2)[:, keep]
        flat = (np.arange(k)[:, None] * nr + idx[None, :]).ravel()
        U = np.bincount(flat, up.ravel(), k * nr).reshape(k, nr)
        L1 = np.bincount(flat, l1.ravel(), k * nr).reshape(k, nr)
        L2 = np.bincount(flat, l2.ravel(), k * nr).reshape(k, nr)
        den = np.sqrt(L1) * np.sqrt(L2)
        out[s0:s0 + k] = np.where(den != 0, U / np.where(den != 0, den, 1), 0.0)
    out[:, 0] = 1.0
    return out
