"""Minimal MRC (MRC2000 / IMOD layout) writer and reader for tilt series and
reconstructed slices, used to hand synthetic data to the command line.

The 1024-byte header layout is the one ICON-GPU reads (mrcfile3.h:52-126 in the
reference); volumes are stored x fastest, then y, then z (sections).  A tilt
series is nx (detector x) x ny (slices, the tilt axis) x nz (tilts).
"""
import struct

import numpy as np

_MODES = {0: np.uint8, 1: np.int16, 2: np.float32, 6: np.uint16}


def header(nx, ny, nz, mode=2, next_bytes=0, amin=0.0, amax=0.0, amean=0.0, rms=0.0):
    """A 1024-byte header: dims, mode, cell = dims, axes 1,2,3, 'MAP ' stamp."""
    h = bytearray(1024)
    struct.pack_into("<10i", h, 0, nx, ny, nz, mode, 0, 0, 0, nx, ny, nz)
    struct.pack_into("<6f", h, 40, float(nx), float(ny), float(nz), 90.0, 90.0, 90.0)
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    struct.pack_into("<3f", h, 76, amin, amax, amean)
    struct.pack_into("<hh", h, 88, 1, 0)
    struct.pack_into("<i", h, 92, next_bytes)
    struct.pack_into("<ii", h, 152, 1146047817, 0)
    h[208:212] = b"MAP "
    h[212] = ord("D")
    struct.pack_into("<f", h, 216, rms)
    return bytes(h)


def write(path, vol, mode=2, next_bytes=0):
    """vol[nz][ny][nx] -> MRC file (converted to the mode's element type)."""
    vol = np.asarray(vol)
    nz, ny, nx = vol.shape
    data = vol.astype(_MODES[mode])
    v = vol.astype(np.float64)
    with open(path, "wb") as fh:
        fh.write(header(nx, ny, nz, mode, next_bytes, float(v.min()), float(v.max()), float(v.mean()),
                        float(v.std())))
        fh.write(b"\0" * next_bytes)
        fh.write(np.ascontiguousarray(data).tobytes())


def write_tilt_series(path, sinos, mode=2):
    """sinos[slices][tilts][nx] -> tilt series nx x slices x tilts."""
    write(path, np.transpose(np.asarray(sinos), (1, 0, 2)), mode)


def read(path):
    """Returns (header fields dict, vol[nz][ny][nx] float32)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    nx, ny, nz, mode = struct.unpack_from("<4i", raw, 0)
    amin, amax, amean = struct.unpack_from("<3f", raw, 76)
    nxt = struct.unpack_from("<i", raw, 92)[0]
    rms = struct.unpack_from("<f", raw, 216)[0]
    mx, my, mz = struct.unpack_from("<3i", raw, 28)
    xlen, ylen, zlen = struct.unpack_from("<3f", raw, 40)
    dt = np.dtype(_MODES[mode])
    nbytes = nx * ny * nz * dt.itemsize
    data = np.frombuffer(raw[len(raw) - nbytes:], dtype=dt).astype(np.float32).reshape(nz, ny, nx)
    return dict(nx=nx, ny=ny, nz=nz, mode=mode, amin=amin, amax=amax, amean=amean, rms=rms, next=nxt,
                mx=mx, my=my, mz=mz, xlen=xlen, ylen=ylen, zlen=zlen), data
