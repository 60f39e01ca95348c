"""MI355X-native ICON reconstruction core.

Python mirror of the C ABI in ``include/icon_mi355x.h`` (ctypes over
``libicon_mi355x.so``).  The compute path is the HIP library; nothing here
computes on the CPU, and loading fails loudly when the library is missing.
"""
from .icon import (  # noqa: F401
    IconError,
    Plan,
    Batch,
    DeviceArray,
    SolveOptions,
    lib,
    lib_path,
    library_loaded,
    fft_rows,
    fft_2d,
    lib_2d2,
    read_tilts,
    partition_slices,
    device_count,
)
