"""Exact (numpy, float64) references shared by the oracle and GPU tests:
the non-uniform DFT on ICON's radial nodes, the centred DFT of projection
rows, and noiseless data that an image explains exactly.  Independent of the
oracle and of the HIP library."""
import numpy as np


def nodes_np(N, tilt):
    """Radial nodes (x, y) of tilt-major node j = i N + k in float64: the line
    f = -0.5 + k/N through the origin at the negated tilt angle."""
    f = -0.5 + np.arange(N) / N
    th = np.deg2rad(-np.asarray(tilt, np.float64))
    return np.stack([np.outer(np.cos(th), f).ravel(), np.outer(np.sin(th), f).ravel()], 1)


def ndft_np(m, xy):
    """Exact transform of the image m [N0][N1] (centred coordinates) at the nodes."""
    N0, N1 = m.shape
    ex = np.exp(-2j * np.pi * np.outer(xy[:, 0], np.arange(N0) - N0 // 2))
    ey = np.exp(-2j * np.pi * np.outer(xy[:, 1], np.arange(N1) - N1 // 2))
    return np.einsum("jx,xy,jy->j", ex, np.asarray(m, np.float64), ey)


def centred_dft_rows(sino):
    """b of every projection row: sum_X p_X exp(-2 pi i (X - N/2)(k - N/2) / N)."""
    N = sino.shape[-1]
    X = np.arange(N) - N // 2
    return (np.asarray(sino, np.float64) @ np.exp(-2j * np.pi * np.outer(X, X) / N).T).ravel()


def consistent_case(N, tilt, seed=11):
    """An image m* whose exact transform is real at every node (point
    symmetric about the centre, first row and column zero) and the projection
    rows whose centred DFT is that transform: noiseless data with A m* = b."""
    rng = np.random.default_rng(seed)
    ms = np.zeros((N, N))
    h = rng.standard_normal((N - 1, N - 1))
    ms[1:, 1:] = h + h[::-1, ::-1]
    bs = ndft_np(ms, nodes_np(N, tilt)).reshape(len(tilt), N)
    X = np.arange(N) - N // 2
    Dinv = np.linalg.inv(np.exp(-2j * np.pi * np.outer(X, X) / N))
    p = bs @ Dinv.T
    assert np.abs(p.imag).max() < 1e-9 * np.abs(p.real).max()
    return ms.astype(np.float32), p.real.astype(np.float32)
