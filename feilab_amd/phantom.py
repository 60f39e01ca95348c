"""Seeded synthetic tilt series of Gaussian-blob phantoms.

Projections are analytic line integrals (no inverse crime): a 2D Gaussian of
width s and amplitude a projects to a 1D Gaussian of amplitude a*sqrt(2 pi)*s
centred on x*cos(t) + z*sin(t).  The geometry follows ICON's convention: t is
the negated tilt-file angle (MainCode.cu:42), x the width coordinate and z the
depth coordinate, both centred at N/2 (nufft_gpu_v8.cu:754-809), so the
Fourier-slice data of a projection row equals the phantom's transform on the
radial nodes of px2D.cu:8-47.

Noise: additive Gaussian with variance var(projection)/SNR per tilt.
"""
import numpy as np

SQRT2PI = np.sqrt(2.0 * np.pi)


def tilt_series(max_deg, step_deg):
    n = int(round(2 * max_deg / step_deg)) + 1
    return np.linspace(-max_deg, max_deg, n).astype(np.float32)


def blobs_2d(N, nblobs, rng, sigma=(2.0, 8.0), amp=(0.2, 1.0), radius=0.4):
    """nblobs isotropic Gaussians in an N x N slice: centres uniform in a disc of
    radius radius*N (centred coordinates), sigma ~ U(sigma), amplitude ~ U(amp)."""
    r = radius * N * np.sqrt(rng.uniform(0, 1, nblobs))
    phi = rng.uniform(0, 2 * np.pi, nblobs)
    return dict(x=r * np.cos(phi), z=r * np.sin(phi), s=rng.uniform(*sigma, nblobs), a=rng.uniform(*amp, nblobs))


def blobs_3d(N, ny, depth, nblobs, rng, sigma=(2.0, 8.0), amp=(0.2, 1.0)):
    """nblobs Gaussians in an N (x) x depth (z) x ny (y) box, centres uniform in the
    inner 80 % of x and z and anywhere in y."""
    return dict(x=rng.uniform(-0.4, 0.4, nblobs) * N, z=rng.uniform(-0.4, 0.4, nblobs) * depth,
                y=rng.uniform(0, ny, nblobs), s=rng.uniform(*sigma, nblobs), a=rng.uniform(*amp, nblobs))


def render_2d(b, N0, N1):
    """The phantom sampled on the reconstruction grid [N0 (x)][N1 (z)]."""
    X = np.arange(N0)[:, None] - N0 // 2
    Z = np.arange(N1)[None, :] - N1 // 2
    img = np.zeros((N0, N1))
    for x, z, s, a in zip(b["x"], b["z"], b["s"], b["a"]):
        img += a * np.exp(-((X - x) ** 2 + (Z - z) ** 2) / (2 * s * s))
    return img.astype(np.float32)


def project_2d(b, N0, tilt_deg, weights=None):
    """Sinogram [nz][N0] of a 2D blob set at the tilt-file angles."""
    th = -np.deg2rad(np.asarray(tilt_deg, np.float64))
    sdet = np.arange(N0) - N0 // 2
    out = np.zeros((len(th), N0))
    w = np.ones(len(b["x"])) if weights is None else weights
    for i, t in enumerate(th):
        c = b["x"] * np.cos(t) + b["z"] * np.sin(t)
        amp = w * b["a"] * SQRT2PI * b["s"]
        out[i] = (amp[None, :] * np.exp(-(sdet[:, None] - c[None, :]) ** 2 / (2 * b["s"][None, :] ** 2))).sum(1)
    return out


def project_3d(b, N0, slices, tilt_deg):
    """Sinograms [len(slices)][nz][N0] of a 3D blob set (tilt axis = y)."""
    th = -np.deg2rad(np.asarray(tilt_deg, np.float64))
    sdet = np.arange(N0) - N0 // 2
    out = np.zeros((len(slices), len(th), N0))
    cos, sin = np.cos(th), np.sin(th)
    slices = np.asarray(slices)
    for x, z, y, s, a in zip(b["x"], b["z"], b["y"], b["s"], b["a"]):
        wy = np.exp(-(slices - y) ** 2 / (2 * s * s))
        keep = wy > 1e-7
        if not keep.any():
            continue
        c = x * cos + z * sin                                  # [nz]
        lo = max(int(np.floor(c.min() + N0 // 2 - 6 * s)), 0)
        hi = min(int(np.ceil(c.max() + N0 // 2 + 6 * s)) + 1, N0)
        if hi <= lo:
            continue
        prof = a * SQRT2PI * s * np.exp(-(sdet[None, lo:hi] - c[:, None]) ** 2 / (2 * s * s))  # [nz][w]
        out[keep, :, lo:hi] += wy[keep, None, None] * prof[None]
    return out


def add_noise(sino, snr, rng):
    """Additive Gaussian noise, variance var(projection)/snr per tilt row."""
    sino = np.asarray(sino, np.float64)
    v = sino.var(axis=-1, keepdims=True)
    return (sino + rng.standard_normal(sino.shape) * np.sqrt(v / snr)).astype(np.float32)


def config1(seed=1001):
    """One 256^2 slice, 61 tilts +-60 step 2, 50 blobs, SNR 0.1."""
    rng = np.random.default_rng(seed)
    tilt = tilt_series(60, 2)
    b = blobs_2d(256, 50, rng)
    clean = project_2d(b, 256, tilt)
    return dict(N=256, tilt=tilt, truth=render_2d(b, 256, 256), clean=clean.astype(np.float32),
                sino=add_noise(clean, 0.1, rng), pre=10, icon=80, post=10)


def config2(nslices=256, seed=1002, N=512, depth=512, nblobs=400, snr=0.05):
    """Tilt series of a 512 x 512 x 256 blob phantom at 41 tilts +-60 step 3, SNR 0.05:
    sinograms of every slice [nslices][41][512]."""
    rng = np.random.default_rng(seed)
    tilt = tilt_series(60, 3)
    b = blobs_3d(N, nslices, depth, nblobs, rng)
    clean = project_3d(b, N, np.arange(nslices), tilt)
    return dict(N=N, tilt=tilt, blobs=b, clean=clean.astype(np.float32), sino=add_noise(clean, snr, rng),
                pre=20, icon=160, post=20)


def render_slices_3d(b, N0, N1, slices):
    """Ground truth of a 3D blob set on the reconstruction grid of each slice:
    [len(slices)][N0 (x)][N1 (z)], the coordinates of project_3d."""
    slices = np.asarray(slices)
    out = np.zeros((len(slices), N0, N1))
    for x, z, y, s, a in zip(b["x"], b["z"], b["y"], b["s"], b["a"]):
        wy = a * np.exp(-(slices - y) ** 2 / (2 * s * s))
        keep = np.nonzero(wy > 1e-7)[0]
        if keep.size == 0:
            continue
        r = int(np.ceil(5 * s))
        cx, cz = x + N0 // 2, z + N1 // 2
        x0, x1 = max(int(np.floor(cx)) - r, 0), min(int(np.ceil(cx)) + r + 1, N0)
        z0, z1 = max(int(np.floor(cz)) - r, 0), min(int(np.ceil(cz)) + r + 1, N1)
        if x1 <= x0 or z1 <= z0:
            continue
        gx = np.exp(-(np.arange(x0, x1) - cx) ** 2 / (2 * s * s))
        gz = np.exp(-(np.arange(z0, z1) - cz) ** 2 / (2 * s * s))
        out[keep, x0:x1, z0:z1] += wy[keep, None, None] * (gx[:, None] * gz[None, :])[None]
    return out.astype(np.float32)


def config2_sino(nslices_seed):
    """Sinograms of config2(nslices, seed) for (nslices, seed): the unit of work
    of a process pool generating training series."""
    n, seed = nslices_seed
    return config2(nslices=n, seed=seed)["sino"]


def config3_specs(seed=1003, nseries=48):
    """The uneven batch: nseries tilt series with width U{128..1024} (the
    thickness equals the width), slice count U{32..300}, maximum tilt
    U(45, 70) deg, tilt step U{1..4} deg and SNR U(0.02, 0.2), seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nseries):
        N = int(rng.integers(128, 1025))
        S = int(rng.integers(32, 301))
        mt = float(rng.uniform(45.0, 70.0))
        st = int(rng.integers(1, 5))
        snr = float(rng.uniform(0.02, 0.2))
        out.append(dict(N=N, S=S, max_tilt=mt, step=st, snr=snr, tilt=tilt_series(mt, st), seed=seed * 100 + i))
    return out


def config3_sino(spec, slices=None):
    """Sinograms [S][nz][N] of one config-3 series: a blob phantom of N x N x S
    (blob count scaled with the area, 400 at 512^2) projected analytically at
    the series' tilts, noise at its SNR.  slices: the first `slices` only."""
    N, S = spec["N"], spec["S"] if slices is None else min(slices, spec["S"])
    rng = np.random.default_rng(spec["seed"])
    b = blobs_3d(N, S, N, max(20, int(400 * (N / 512.0) ** 2)), rng)
    return add_noise(project_3d(b, N, np.arange(S), spec["tilt"]), spec["snr"], rng)
