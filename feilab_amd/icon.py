"""ctypes mirror of the MI355X ICON C ABI (include/icon_mi355x.h).

The reference's in-process interface is ``Lib_2D2`` plus the ``cunfft_*`` plan
API (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/MainCode.cuh:10,
nufft_gpu_v8.cuh:153-223).  This module exposes the same operations with the
same argument meaning over the HIP library:

* :class:`Plan` -- ``cunfft_initPlan_2d2`` + ``_1d`` + ``initGICONPara2`` for one
  tilt series on one device, with batched forward / adjoint NUFFTs, the 1D
  Fourier-slice transform of the projections and the ICON solve.
* :func:`lib_2d2` -- ``Lib_2D2``: file-level reconstruction of a slice range.

Arrays cross the boundary as device buffers (:class:`DeviceArray`) or as host
numpy arrays that are copied in and out.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libicon_mi355x.so")
_lib = None

ICON_OK = 0
_STATUS = {1: "ICON_ERR_ARG", 2: "ICON_ERR_DEVICE", 3: "ICON_ERR_IO", 4: "ICON_ERR_CAPACITY", 5: "ICON_ERR_STATE"}


class IconError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_STATUS.get(code, code), msg))
        self.code = code


class PlanInfo(ctypes.Structure):
    _fields_ = [("N0", ctypes.c_int32), ("N1", ctypes.c_int32), ("nz", ctypes.c_int32), ("M", ctypes.c_int32),
                ("n0", ctypes.c_int32), ("n1", ctypes.c_int32), ("m", ctypes.c_int32), ("iomit", ctypes.c_int32),
                ("sigma0", ctypes.c_float), ("sigma1", ctypes.c_float), ("b0", ctypes.c_float),
                ("b1", ctypes.c_float), ("max_slices", ctypes.c_int32), ("col_block", ctypes.c_int32),
                ("window", ctypes.c_int32), ("device_bytes", ctypes.c_int64)]


# icon_window
WINDOWS = {"gauss": 0, "kb6": 1}


class SolveOptions(ctypes.Structure):
    """The reference's before[] / betain[] (ICONGPU.cu:380-387)."""
    _fields_ = [("pre", ctypes.c_int32), ("icon", ctypes.c_int32), ("post", ctypes.c_int32),
                ("data_type", ctypes.c_int32), ("threshold", ctypes.c_float), ("use_graph", ctypes.c_int32),
                ("reuse_data", ctypes.c_int32)]


def lib_path():
    return _LIB


def library_loaded():
    return _lib is not None


def lib():
    """Loads libicon_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise ImportError("libicon_mi355x.so is not built; run __graft_entry__.build() "
                          "(make -C feilab_amd/csrc)")
    L = ctypes.CDLL(_LIB)
    vp, ci, cf, cs = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    L.icon_last_error.restype = ctypes.c_char_p
    L.icon_version.restype = ctypes.c_char_p
    L.icon_lib_2d2_error.restype = ctypes.c_char_p
    sig = {
        "icon_device_count": [ctypes.POINTER(ci)],
        "icon_device_compute_units": [ci, ctypes.POINTER(ci)],
        "icon_malloc": [ctypes.POINTER(vp), cs],
        "icon_free": [vp],
        "icon_memcpy_h2d": [vp, vp, cs],
        "icon_memcpy_d2h": [vp, vp, cs],
        "icon_memset_d": [vp, ci, cs],
        "icon_memcpy_d2d_async": [vp, vp, cs, vp],
        "icon_stream_create": [ctypes.POINTER(vp)],
        "icon_stream_destroy": [vp],
        "icon_stream_synchronize": [vp],
        "icon_event_create": [ctypes.POINTER(vp)],
        "icon_event_destroy": [vp],
        "icon_event_record": [vp, vp],
        "icon_stream_wait_event": [vp, vp],
        "icon_device_synchronize": [],
        "icon_plan_create": [ctypes.POINTER(vp), ci, ci, ci, ci, ctypes.POINTER(cf), ci],
        "icon_plan_create_cv": [ctypes.POINTER(vp), ci, ci, ci, ci, ctypes.POINTER(cf), ci],
        "icon_plan_create_window": [ctypes.POINTER(vp), ci, ci, ci, ci, ctypes.POINTER(cf), ci, ci, ci],
        "icon_plan_destroy": [vp],
        "icon_plan_get_info": [vp, ctypes.POINTER(PlanInfo)],
        "icon_plan_set_col_block": [vp, ci],
        "icon_plan_autotune": [vp, ci, ctypes.POINTER(ci), ctypes.POINTER(ctypes.c_double)],
        "icon_nufft_forward_real": [vp, ci, vp, vp, vp],
        "icon_nufft_forward_complex": [vp, ci, vp, vp, vp],
        "icon_nufft_adjoint": [vp, ci, vp, vp, vp],
        "icon_sino_to_b": [vp, ci, vp, vp, vp],
        "icon_fft_rows": [ci, ci, ci, vp, vp],
        "icon_fft_2d": [ci, ci, ci, ci, vp, vp],
        "icon_reconstruct": [vp, ci, vp, vp, vp, vp, vp, vp, ctypes.POINTER(SolveOptions), vp],
        "icon_reconstruct_host": [vp, ci, vp, vp, vp, vp, vp, vp, ctypes.POINTER(SolveOptions)],
        "icon_lib_2d2": [ci, ci, ci, ci, ctypes.c_char_p, vp, ctypes.c_char_p, ctypes.c_char_p,
                         ctypes.POINTER(ci), ctypes.POINTER(cf), ci],
        "icon_lib_2d_crossv": [ci, ci, ci, ci, ctypes.c_char_p, vp, ctypes.c_char_p, ctypes.c_char_p,
                               ctypes.POINTER(ci), ctypes.POINTER(cf), ci],
        "icon_frc": [vp, vp, ci, ci, vp],
        "icon_preprocess": [ci, vp, ci, ci, ci, vp, ci, vp, vp],
        "icon_preprocess_noalloc": [ci, vp, ci, ci, ci, vp, ci, vp, vp],
        "icon_preprocess_file": [ci, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ci],
        "icon_mask3_weights": [ci, vp, ci, cf, cf, ci, ci, cf, vp],
        "icon_mask3_filter": [ci, vp, ci, ci, vp, vp],
        "icon_mask3": [ci, ctypes.c_char_p, ci, ci, ci, ctypes.c_char_p, ctypes.c_char_p, ci, ctypes.c_char_p,
                       ctypes.c_char_p, ci, ci, ci],
        "icon_read_tilts": [ctypes.c_char_p, ctypes.POINTER(cf), ci, ctypes.POINTER(ci)],
        "icon_partition_slices": [ci, ctypes.POINTER(ci), ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ci)],
        "icon_geometry_tables": [ci, ci, ci, ctypes.POINTER(cf), ctypes.POINTER(PlanInfo)] + [vp] * 10,
        "icon_mrc_read_y_slices": [ctypes.c_char_p, ci, ci, vp, ctypes.POINTER(ci)],
        "icon_mrc_write_slice": [ctypes.c_char_p, ctypes.c_char_p, vp, ci, ci],
        "icon_time_kernel": [vp, ci, ci, ci, ctypes.POINTER(ctypes.c_double)],
        "icon_copy_probe": [vp, vp, cs, ci, ctypes.POINTER(ctypes.c_double)],
        "icon_reconstruct_fp64": [ci, ci, ci, ci, ctypes.POINTER(cf), ci, vp, vp, vp, vp, vp,
                                  ctypes.POINTER(SolveOptions)],
        "icon_batch_create": [ctypes.POINTER(vp), ci, vp, vp],
        "icon_batch_destroy": [vp],
        "icon_batch_reconstruct": [vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(SolveOptions), vp],
        "icon_batch_launches_per_iteration": [vp, ctypes.POINTER(ci)],
        "icon_batch_autotune": [vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ci
    L.icon_preprocess_error.restype = ctypes.c_char_p
    L.icon_mask3_error.restype = ctypes.c_char_p
    L.icon_frc_smooth.argtypes = [vp, ci, ci]
    L.icon_frc_smooth.restype = None
    _lib = L
    return L


def _check(rc, what=""):
    if rc != ICON_OK:
        msg = lib().icon_last_error().decode(errors="replace")
        raise IconError(rc, (what + ": " if what else "") + msg)


def device_count():
    n = ctypes.c_int(0)
    _check(lib().icon_device_count(ctypes.byref(n)), "icon_device_count")
    return n.value


class DeviceArray:
    """A device buffer owned by Python (icon_malloc / icon_free)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        _check(lib().icon_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "icon_malloc")

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _check(lib().icon_memcpy_h2d(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "icon_memcpy_h2d")
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().icon_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "icon_memcpy_h2d")

    def zero(self):
        _check(lib().icon_memset_d(self.ptr, 0, self.nbytes), "icon_memset_d")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(lib().icon_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes),
                   "icon_memcpy_d2h")
        return out

    def addr(self, offset_bytes=0):
        """Device address offset_bytes into the buffer (for copies of parts)."""
        if not 0 <= offset_bytes <= self.nbytes:
            raise ValueError("offset %d outside a buffer of %d bytes" % (offset_bytes, self.nbytes))
        return self.ptr.value + offset_bytes

    def free(self):
        if self.ptr and self.ptr.value and _lib is not None:
            _lib.icon_free(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream of the current device (default priority)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        _check(lib().icon_stream_create(ctypes.byref(self.h)), "icon_stream_create")

    @property
    def value(self):
        return self.h.value

    def synchronize(self):
        _check(lib().icon_stream_synchronize(self.h), "icon_stream_synchronize")

    def record(self, event):
        _check(lib().icon_event_record(event.h, self.h), "icon_event_record")

    def wait(self, event):
        _check(lib().icon_stream_wait_event(self.h, event.h), "icon_stream_wait_event")

    def close(self):
        if self.h and self.h.value and _lib is not None:
            _lib.icon_stream_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _check(lib().icon_event_create(ctypes.byref(self.h)), "icon_event_create")

    def close(self):
        if self.h and self.h.value and _lib is not None:
            _lib.icon_event_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_d2d(dst, src, nbytes, stream=None):
    """Asynchronous device-to-device copy between device addresses (ints or DeviceArrays)."""
    _check(lib().icon_memcpy_d2d_async(_ptr(dst), _ptr(src), int(nbytes), _ptr(stream)), "icon_memcpy_d2d_async")


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, Stream):
        return x.h
    return ctypes.c_void_p(int(x))


class Plan:
    """NUFFT/ICON plan of one tilt series on one device.

    ``tilt_deg`` are the angles as written in the tilt file; the plan negates
    them (MainCode.cu:42).  Images are [S][nx][thickness] (index X*N1 + Y),
    node vectors [S][nz*nx] complex64, sinograms [S][nz][nx] float32.
    """

    def __init__(self, nx, tilt_deg, thickness=0, max_slices=1, device=0, cross_validation=False,
                 window="gauss"):
        """cross_validation=True: the plan of Lib_2D_crossV2 (the least-|angle|
        tilt left out of the data and used as the reprojection angle).
        window: "gauss" (the reference's 14-tap Gaussian) or "kb6" (6-tap
        Kaiser-Bessel)."""
        tilt = np.ascontiguousarray(tilt_deg, np.float32)
        self.tilt, self.thickness, self.window, self.device = tilt.copy(), int(thickness), window, device
        self.h = ctypes.c_void_p()
        tp = tilt.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if window == "gauss":
            create = lib().icon_plan_create_cv if cross_validation else lib().icon_plan_create
            rc = create(ctypes.byref(self.h), device, int(nx), int(thickness), len(tilt), tp, int(max_slices))
        else:
            rc = lib().icon_plan_create_window(ctypes.byref(self.h), device, int(nx), int(thickness), len(tilt), tp,
                                               int(max_slices), int(bool(cross_validation)), WINDOWS[window])
        _check(rc, "icon_plan_create")
        self.cross_validation = cross_validation
        info = PlanInfo()
        _check(lib().icon_plan_get_info(self.h, ctypes.byref(info)), "icon_plan_get_info")
        self.info = info
        self.N0, self.N1, self.nz, self.M = info.N0, info.N1, info.nz, info.M
        self.n0, self.n1, self.iomit = info.n0, info.n1, info.iomit
        self.max_slices = info.max_slices

    def close(self):
        if self.h and self.h.value and _lib is not None:
            _lib.icon_plan_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_col_block(self, c):
        _check(lib().icon_plan_set_col_block(self.h, int(c)), "icon_plan_set_col_block")
        self.info.col_block = int(c)

    def autotune(self, S):
        """Picks the fastest column tile for S slices; returns (C, {C: ms per iteration's column kernels})."""
        c = ctypes.c_int(0)
        ms = (ctypes.c_double * 2)()
        _check(lib().icon_plan_autotune(self.h, int(S), ctypes.byref(c), ms), "icon_plan_autotune")
        self.info.col_block = c.value
        return c.value, {w: ms[i] for i, w in enumerate((8, 16)) if ms[i] > 0}

    # -- batched operators (device in / device out) -------------------------
    def forward_real_dev(self, S, d_img, d_f, stream=None):
        _check(lib().icon_nufft_forward_real(self.h, S, _ptr(d_img), _ptr(d_f), _ptr(stream)), "forward_real")

    def forward_complex_dev(self, S, d_img, d_f, stream=None):
        _check(lib().icon_nufft_forward_complex(self.h, S, _ptr(d_img), _ptr(d_f), _ptr(stream)), "forward_complex")

    def adjoint_dev(self, S, d_f, d_img, stream=None):
        _check(lib().icon_nufft_adjoint(self.h, S, _ptr(d_f), _ptr(d_img), _ptr(stream)), "adjoint")

    def sino_to_b_dev(self, S, d_sino, d_b, stream=None):
        _check(lib().icon_sino_to_b(self.h, S, _ptr(d_sino), _ptr(d_b), _ptr(stream)), "sino_to_b")

    def reconstruct_dev(self, S, d_sino, d_m, opts, d_m_init=None, d_reproj=None, d_resid=None, d_alpha=None,
                        stream=None):
        _check(lib().icon_reconstruct(self.h, S, _ptr(d_sino), _ptr(d_m), _ptr(d_m_init), _ptr(d_reproj),
                                      _ptr(d_resid), _ptr(d_alpha), ctypes.byref(opts), _ptr(stream)),
               "icon_reconstruct")

    KERNELS = {"row_fwd": 0, "col_fwd": 1, "combine": 2, "col_adj": 3, "row_adj": 4, "sino": 5,
               "combine_alpha": 6}

    def time_kernel(self, name, S, reps=20):
        """Average launch time (ms) of one iteration kernel over S slices (HIP events on the plan stream)."""
        ms = ctypes.c_double(0)
        _check(lib().icon_time_kernel(self.h, self.KERNELS[name], int(S), int(reps), ctypes.byref(ms)),
               "icon_time_kernel")
        return ms.value

    # -- host conveniences ---------------------------------------------------
    def forward(self, img):
        img = np.asarray(img)
        S = img.shape[0] if img.ndim == 3 else 1
        cplx = np.iscomplexobj(img)
        d_img = DeviceArray.from_host(img.reshape(S, self.N0, self.N1), np.complex64 if cplx else np.float32)
        d_f = DeviceArray((S, self.M), np.complex64)
        (self.forward_complex_dev if cplx else self.forward_real_dev)(S, d_img, d_f)
        out = d_f.numpy()
        return out if img.ndim == 3 else out[0]

    def adjoint(self, f):
        f = np.asarray(f)
        S = f.shape[0] if f.ndim == 2 else 1
        d_f = DeviceArray.from_host(f.reshape(S, self.M), np.complex64)
        d_img = DeviceArray((S, self.N0, self.N1), np.complex64)
        self.adjoint_dev(S, d_f, d_img)
        out = d_img.numpy()
        return out if f.ndim == 2 else out[0]

    def sino_to_b(self, sino):
        sino = np.asarray(sino, np.float32)
        S = sino.shape[0] if sino.ndim == 3 else 1
        d_s = DeviceArray.from_host(sino.reshape(S, self.nz, self.N0), np.float32)
        d_b = DeviceArray((S, self.M), np.complex64)
        self.sino_to_b_dev(S, d_s, d_b)
        out = d_b.numpy()
        return out if sino.ndim == 3 else out[0]

    def solve(self, sino, pre, icon, post, data_type=2, threshold=0.0, m_init=None, trace=False,
              reproject=False, use_graph=False):
        """ICON solve of one or more slices (host arrays in and out)."""
        sino = np.ascontiguousarray(sino, np.float32)
        single = sino.ndim == 2
        S = 1 if single else sino.shape[0]
        iters = pre + icon + post
        opts = SolveOptions(pre, icon, post, data_type, threshold, 1 if use_graph else 0)
        m = np.zeros((S, self.N0, self.N1), np.float32)
        mi = None if m_init is None else np.ascontiguousarray(m_init, np.float32).reshape(S, self.N0, self.N1)
        rp = np.zeros((S, self.N0), np.float32) if reproject else None
        rs = np.zeros((S, iters + 1)) if trace else None
        al = np.zeros((S, max(iters, 1))) if trace else None
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
        _check(lib().icon_reconstruct_host(self.h, S, P(sino), P(m), P(mi), P(rp), P(rs), P(al),
                                           ctypes.byref(opts)), "icon_reconstruct_host")
        out = [m[0] if single else m]
        if reproject:
            out.append(rp[0] if single else rp)
        if trace:
            out.append(rs[0] if single else rs)
            out.append((al[0] if single else al)[..., :iters])
        return out[0] if len(out) == 1 else tuple(out)


def reconstruct_fp64(nx, tilt_deg, sino, pre, icon, post, data_type=2, threshold=0.0, thickness=0, device=0):
    """The all-FP64 device solve (icon_reconstruct_fp64) of sinograms [S][nz][nx]:
    (images [S][nx][thickness], reprojections [S][nx], residual traces
    [S][iters+1], step sizes [S][iters]), all float64."""
    tilt = np.ascontiguousarray(tilt_deg, np.float32)
    sino = np.ascontiguousarray(sino, np.float32)
    if sino.ndim == 2:
        sino = sino[None]
    S, nz = sino.shape[0], len(tilt)
    N1 = thickness or nx
    iters = pre + icon + post
    m = np.zeros((S, nx, N1))
    rp = np.zeros((S, nx))
    rs = np.zeros((S, iters + 1))
    al = np.zeros((S, max(iters, 1)))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _check(lib().icon_reconstruct_fp64(device, int(nx), int(thickness), nz,
                                       tilt.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), S, P(sino), P(m), P(rp),
                                       P(rs), P(al), ctypes.byref(SolveOptions(pre, icon, post, data_type, threshold,
                                                                               0, 0))), "icon_reconstruct_fp64")
    return m, rp, rs, al[:, :iters]


class Batch:
    """Several series (one Plan each, same device) solved in the same packed
    launches: icon_batch_create / icon_batch_reconstruct.  slices[i] slices of
    series i are solved by every call."""

    def __init__(self, plans, slices):
        self.plans = list(plans)
        self.slices = [int(v) for v in slices]
        n = len(self.plans)
        arr = (ctypes.c_void_p * n)(*[p.h.value for p in self.plans])
        sl = (ctypes.c_int * n)(*self.slices)
        self.h = ctypes.c_void_p()
        _check(lib().icon_batch_create(ctypes.byref(self.h), n, arr, sl), "icon_batch_create")

    def close(self):
        if self.h and self.h.value and _lib is not None:
            _lib.icon_batch_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arr(xs, n):
        if xs is None:
            return None
        return (ctypes.c_void_p * n)(*[None if x is None else (x.ptr.value if isinstance(x, DeviceArray) else int(x))
                                       for x in xs])

    def reconstruct_dev(self, d_sinos, d_ms, opts, d_m_inits=None, d_reprojs=None, d_resids=None, d_alphas=None,
                        stream=None):
        n = len(self.plans)
        A = lambda xs: self._arr(xs, n)
        _check(lib().icon_batch_reconstruct(self.h, A(d_sinos), A(d_ms), A(d_m_inits), A(d_reprojs), A(d_resids),
                                            A(d_alphas), ctypes.byref(opts), _ptr(stream)), "icon_batch_reconstruct")

    def launches_per_iteration(self):
        v = ctypes.c_int(0)
        _check(lib().icon_batch_launches_per_iteration(self.h, ctypes.byref(v)), "icon_batch_launches_per_iteration")
        return v.value

    def autotune(self):
        """Column tile per size class; returns the tile chosen for each series."""
        out = (ctypes.c_int * len(self.plans))()
        _check(lib().icon_batch_autotune(self.h, out), "icon_batch_autotune")
        for p, c in zip(self.plans, out):
            p.info.col_block = c
        return list(out)

    def solve(self, sinos, pre, icon, post, data_type=2, threshold=0.0, trace=False, reproject=False):
        """Host arrays in and out: sinos[i] [slices[i]][nz_i][N0_i] -> images
        (and residual traces [S][iters+1], reprojections [S][N0]) per series."""
        iters = pre + icon + post
        ds, dm, dr, dp = [], [], [], []
        for p, S, x in zip(self.plans, self.slices, sinos):
            ds.append(DeviceArray.from_host(np.asarray(x, np.float32).reshape(S, p.nz, p.N0)))
            dm.append(DeviceArray((S, p.N0, p.N1), np.float32))
            dr.append(DeviceArray((S, iters + 1), np.float64) if trace else None)
            dp.append(DeviceArray((S, p.N0), np.float32) if reproject else None)
        self.reconstruct_dev(ds, dm, SolveOptions(pre, icon, post, data_type, threshold, 1, 0),
                             d_resids=dr if trace else None, d_reprojs=dp if reproject else None)
        lib().icon_device_synchronize()
        out = [d.numpy() for d in dm]
        res = [d.numpy() for d in dr] if trace else None
        rps = [d.numpy() for d in dp] if reproject else None
        return out, res, rps


def fft_rows(x, sign=-1):
    x = np.ascontiguousarray(x, np.complex64)
    n = x.shape[-1]
    d = DeviceArray.from_host(x)
    _check(lib().icon_fft_rows(n, x.size // n, sign, d.ptr, None), "icon_fft_rows")
    return d.numpy()


def fft_2d(x, sign=-1):
    x = np.ascontiguousarray(x, np.complex64)
    n0, n1 = x.shape[-2:]
    d = DeviceArray.from_host(x)
    _check(lib().icon_fft_2d(n0, n1, x.size // (n0 * n1), sign, d.ptr, None), "icon_fft_2d")
    return d.numpy()


def copy_probe(nbytes=1 << 30, reps=20):
    """HBM bandwidth of a plain device copy: returns (ms per copy, GB/s counting read + write)."""
    src = DeviceArray((nbytes // 4,), np.float32)
    dst = DeviceArray((nbytes // 4,), np.float32)
    ms = ctypes.c_double(0)
    _check(lib().icon_copy_probe(dst.ptr, src.ptr, nbytes, reps, ctypes.byref(ms)), "icon_copy_probe")
    return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9


def geometry_tables(nx, tilt_deg, thickness=0):
    """Host-side plan geometry (no device): the tables icon_plan_create uploads."""
    tilt = np.ascontiguousarray(tilt_deg, np.float32)
    nz = len(tilt)
    info = PlanInfo()
    _check(lib().icon_geometry_tables(nx, thickness, nz, tilt.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      ctypes.byref(info), *([None] * 10)), "icon_geometry_tables")
    N0, N1, M = info.N0, info.N1, info.M
    out = dict(nodes=np.zeros((M, 2), np.float32), w=np.zeros(M, np.float32),
               phix=np.zeros((M, 14), np.float32), phiy=np.zeros((M, 14), np.float32),
               ix0=np.zeros(M, np.int32), iy0=np.zeros(M, np.int32), ck=np.zeros((N0, N1), np.float32),
               ck1=np.zeros(N0, np.float32), phi1=np.zeros((N0, 14), np.float32), i10=np.zeros(N0, np.int32))
    order = ["nodes", "w", "phix", "phiy", "ix0", "iy0", "ck", "ck1", "phi1", "i10"]
    _check(lib().icon_geometry_tables(nx, thickness, nz, tilt.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      ctypes.byref(info), *[out[k].ctypes.data_as(ctypes.c_void_p) for k in order]),
           "icon_geometry_tables")
    out["info"] = info
    return out


def mrc_read_y_slices(path, y0, count):
    """Slices [y0, y0+count) of an MRC tilt series as [count][nz][nx] float32 (mrc_read_slice 'y')."""
    dims = (ctypes.c_int * 4)()
    rc = lib().icon_mrc_read_y_slices(path.encode(), 0, 0, None, dims)
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))
    nx, ny, nz = dims[0], dims[1], dims[2]
    out = np.zeros((count, nz, nx), np.float32)
    rc = lib().icon_mrc_read_y_slices(path.encode(), y0, count, out.ctypes.data_as(ctypes.c_void_p), dims)
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))
    return out


def mrc_write_slice(path, template_path, m):
    """Writes a reconstructed slice m[width][thickness] as ICON's midYYYYY.mrc."""
    m = np.ascontiguousarray(m, np.float32)
    rc = lib().icon_mrc_write_slice(path.encode(), template_path.encode(), m.ctypes.data_as(ctypes.c_void_p),
                                    m.shape[0], m.shape[1])
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))


def read_tilts(path, cap=1000):
    buf = (ctypes.c_float * cap)()
    n = ctypes.c_int(0)
    rc = lib().icon_read_tilts(path.encode(), buf, cap, ctypes.byref(n))
    if rc == 3:
        raise IconError(rc, "Can not open anglefile '%s'!" % path)
    if rc:
        raise IconError(rc, "tilt file holds more than %d values" % cap)
    return np.array(buf[:n.value], np.float32)


def partition_slices(cus, y_start, y_end):
    cus = list(cus)
    nd = len(cus)
    c = (ctypes.c_int * nd)(*cus)
    f = (ctypes.c_int * nd)()
    l_ = (ctypes.c_int * nd)()
    used = lib().icon_partition_slices(nd, c, y_start, y_end, f, l_)
    if used < 0 or used > nd:
        raise IconError(1, "partition failed")
    return [(f[i], l_[i]) for i in range(used)]


def frc(image1, image2):
    """calFRC_nonsqure (calFRC.cu:93-174) of two [ny][nx] images: nx/2 rings."""
    a = np.ascontiguousarray(image1, np.float32)
    b = np.ascontiguousarray(image2, np.float32)
    ny, nx = a.shape
    out = np.zeros(nx // 2 + 1, np.float32)
    rc = lib().icon_frc(a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p), nx, ny,
                        out.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))
    return out[:nx // 2]


def frc_smooth(curve, overlap=10):
    """frcSmooth (ICONGPU.cu:59-91) of an FRC curve (returns a new array)."""
    c = np.ascontiguousarray(curve, np.float32).copy()
    lib().icon_frc_smooth(c.ctypes.data_as(ctypes.c_void_p), len(c), int(overlap))
    return c


def preprocess_dev(d_stack, nz, ny, nx, tilt_deg=None, thickness=0, device=0, noalloc=False):
    """ICONPreProcess in place on a device stack [nz][ny][nx] (a DeviceArray);
    returns the subtracted modes [nz].  noalloc: icon_preprocess_noalloc (no
    device allocation, at most 4096 images)."""
    modes = np.zeros(nz)
    tl = None if tilt_deg is None else np.ascontiguousarray(tilt_deg, np.float32)
    fn = lib().icon_preprocess_noalloc if noalloc else lib().icon_preprocess
    rc = fn(device, d_stack.ptr, nx, ny, nz, None if tl is None else tl.ctypes.data_as(ctypes.c_void_p),
            int(thickness), modes.ctypes.data_as(ctypes.c_void_p), None)
    if rc:
        raise IconError(rc, lib().icon_preprocess_error().decode(errors="replace"))
    return modes


def preprocess(stack, tilt_deg=None, thickness=0, device=0):
    """ICONPreProcess on the device (ICONPreProcess.c:31-86): stack [nz][ny][nx]
    -> (processed stack, subtracted modes [nz])."""
    st = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = st.shape
    d = DeviceArray.from_host(st)
    modes = np.zeros(nz)
    tl = None if tilt_deg is None else np.ascontiguousarray(tilt_deg, np.float32)
    rc = lib().icon_preprocess(device, d.ptr, nx, ny, nz, None if tl is None else tl.ctypes.data_as(ctypes.c_void_p),
                               int(thickness), modes.ctypes.data_as(ctypes.c_void_p), None)
    if rc:
        raise IconError(rc, lib().icon_preprocess_error().decode(errors="replace"))
    return d.numpy(), modes


def preprocess_file(in_path, out_path, tilt_path=None, thickness=0, device=0):
    rc = lib().icon_preprocess_file(device, in_path.encode(), out_path.encode(),
                                    None if tilt_path is None else tilt_path.encode(), int(thickness))
    if rc:
        raise IconError(rc, lib().icon_preprocess_error().decode(errors="replace"))


def mask3_weights(N, tilt_deg, radius05_rate, radius03_rate, radius_given, gauss_width, mean_step):
    """ICONMask3's half-spectrum multiplier [N][N/2+1] (host computation)."""
    t = np.ascontiguousarray(tilt_deg, np.float32)
    out = np.zeros((N, N // 2 + 1), np.float32)
    rc = lib().icon_mask3_weights(N, t.ctypes.data_as(ctypes.c_void_p), len(t), radius05_rate, radius03_rate,
                                  int(radius_given), int(gauss_width), mean_step, out.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise IconError(rc, lib().icon_mask3_error().decode(errors="replace"))
    return out


def mask3_filter(slices, half_mask, device=0):
    """ICONMask3's filtering of [S][N][N] real slices on the device."""
    sl = np.ascontiguousarray(slices, np.float32)
    d = DeviceArray.from_host(sl)
    hm = np.ascontiguousarray(half_mask, np.float32)
    rc = lib().icon_mask3_filter(device, d.ptr, sl.shape[-1], sl.shape[0], hm.ctypes.data_as(ctypes.c_void_p), None)
    if rc:
        raise IconError(rc, lib().icon_mask3_error().decode(errors="replace"))
    return d.numpy()


def lib_2d_crossv(mrc_path, angle_path, result_dir, slice_begin, slice_end, pre, icon, post, data_type=1,
                  threshold=0.03, thickness=0, device=0, batch=0):
    """Lib_2D_crossV2 (MainCode_crossV.cu:3-189): the slices solved without the least-|angle| tilt;
    returns the reprojections at that angle [slices][nx] (no files written)."""
    import struct
    with open(mrc_path, "rb") as fh:
        nx = struct.unpack("<i", fh.read(4))[0]
    rp = np.zeros((max(slice_end - slice_begin, 0), nx), np.float32)
    before = (ctypes.c_int * 5)(pre, post, 1, data_type, 0)
    betain = (ctypes.c_float * 2)(threshold, float(thickness))
    rc = lib().icon_lib_2d_crossv(device, slice_begin, slice_end, pre + icon + post, mrc_path.encode(),
                                  rp.ctypes.data_as(ctypes.c_void_p), angle_path.encode(), result_dir.encode(),
                                  before, betain, batch)
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))
    return rp


def lib_2d2(mrc_path, angle_path, result_dir, slice_begin, slice_end, pre, icon, post, data_type=1,
            threshold=0.03, thickness=0, device=0, batch=0):
    """Lib_2D2 (MainCode.cu:3-177): reconstruct [slice_begin, slice_end) of a tilt series into
    result_dir/midYYYYY.mrc; returns the reprojection rows [slices][nx]."""
    import struct
    with open(mrc_path, "rb") as fh:
        nx = struct.unpack("<i", fh.read(4))[0]
    rp = np.zeros((max(slice_end - slice_begin, 0), nx), np.float32)
    before = (ctypes.c_int * 5)(pre, post, 1, data_type, 1)
    betain = (ctypes.c_float * 2)(threshold, float(thickness))
    rc = lib().icon_lib_2d2(device, slice_begin, slice_end, pre + icon + post, mrc_path.encode(),
                            rp.ctypes.data_as(ctypes.c_void_p), angle_path.encode(), result_dir.encode(), before,
                            betain, batch)
    if rc:
        raise IconError(rc, lib().icon_lib_2d2_error().decode(errors="replace"))
    return rp
