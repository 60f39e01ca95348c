"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of ICON-GPU's reconstruction core (see the
header of icon_oracle.cpp for the reference file:line of every function).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import
this module, and only as the checker or the CPU baseline -- never as the path
that is measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libicon_oracle.so")
_REF_MRC = os.path.join(_HERE, "_ref", "libmrcfile3_ref.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
TAPS = 14


def build(ref=True):
    """Compile the oracle (and the reference MRC library when the reference is mounted)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if ref and os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build(ref=False)
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_grid.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
        L.orc_window_params.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]
        L.orc_weight.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_px2d.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_px1d.argtypes = [_f32p, ctypes.c_int]
        L.orc_read_angles.argtypes = [ctypes.c_char_p, _f32p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.orc_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_plan_create.restype = vp
        L.orc_plan_destroy.argtypes = [vp]
        L.orc_plan_info.argtypes = [vp, _i32p, _f32p]
        L.orc_plan_tables.argtypes = [vp] + [vp] * 10
        L.orc_forward.argtypes = [vp, _f64p, _f64p]
        L.orc_adjoint.argtypes = [vp, _f64p, _f64p]
        L.orc_sino_to_b.argtypes = [vp, _f32p, _f64p]
        L.orc_adjoint1d.argtypes = [vp, _f64p, _f64p]
        L.orc_fft.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_fft2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_ndft_forward.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f64p, _f64p]
        L.orc_ndft_adjoint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f64p, _f64p]
        L.orc_icon_solve.argtypes = [vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, _f32p, vp, vp, vp]
        L.orc_icon_solve.restype = ctypes.c_int
        L.orc_icon_solve_batch.argtypes = [vp, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_float, _f32p, vp, ctypes.c_int]
        L.orc_icon_solve_batch.restype = ctypes.c_int
        L.orc_reproject.argtypes = [vp, _f32p, _f32p]
        L.orc_preprocess.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p]
        L.orc_preprocess.restype = None
        _lib = L
    return _lib


def _c2d(z):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    return z.view(np.float64).reshape(-1)


def _d2c(d, shape):
    return d.view(np.complex128).reshape(shape)


def grid(N):
    s = ctypes.c_float()
    n = ctypes.c_int()
    lib().orc_grid(N, ctypes.byref(s), ctypes.byref(n))
    return s.value, n.value


def window_params(sigma, m=6):
    b = ctypes.c_float()
    bf = ctypes.c_float()
    lib().orc_window_params(sigma, m, ctypes.byref(b), ctypes.byref(bf))
    return b.value, bf.value


def weight(nx, theta):
    theta = np.ascontiguousarray(theta, np.float32)
    w = np.zeros(nx * len(theta), np.float32)
    lib().orc_weight(w, nx, len(theta), theta)
    return w.reshape(len(theta), nx)


def px2d(nx, theta):
    theta = np.ascontiguousarray(theta, np.float32)
    xy = np.zeros(2 * nx * len(theta), np.float32)
    lib().orc_px2d(xy, nx, len(theta), theta)
    return xy.reshape(-1, 2)


def px1d(nx):
    x = np.zeros(nx, np.float32)
    lib().orc_px1d(x, nx)
    return x


def read_angles(path, cap=180):
    th = np.zeros(cap, np.float32)
    om = ctypes.c_float()
    n = lib().orc_read_angles(path.encode(), th, cap, ctypes.byref(om))
    if n < 0:
        raise OSError("cannot open angle file %s" % path)
    return th[:n].copy(), om.value


def fft(x, sign=-1, prec=64):
    d = _c2d(x).copy()
    lib().orc_fft(len(x), sign, prec, d)
    return _d2c(d, (len(x),))


def fft2(x, sign=-1, prec=64):
    n0, n1 = x.shape
    d = _c2d(x).copy()
    lib().orc_fft2(n0, n1, sign, prec, d)
    return _d2c(d, (n0, n1))


def ndft_forward(nodes, img):
    N0, N1 = img.shape
    nodes = np.ascontiguousarray(nodes, np.float32).reshape(-1)
    M = len(nodes) // 2
    f = np.zeros(2 * M)
    lib().orc_ndft_forward(N0, N1, M, nodes, _c2d(img), f)
    return _d2c(f, (M,))


def ndft_adjoint(nodes, f, N0, N1):
    nodes = np.ascontiguousarray(nodes, np.float32).reshape(-1)
    M = len(nodes) // 2
    out = np.zeros(2 * N0 * N1)
    lib().orc_ndft_adjoint(N0, N1, M, nodes, _c2d(f), out)
    return _d2c(out, (N0, N1))


class Plan:
    """One series' NUFFT plan in the oracle (prec 32 = FP32 mirror, 64 = FP64).

    ``tilt_deg`` are the tilt-file angles; they are negated as the reference
    does on read (MainCode.cu:42)."""

    def __init__(self, N0, N1, tilt_deg, prec=32, m=6):
        self.theta = np.ascontiguousarray(-np.asarray(tilt_deg, np.float32), np.float32)
        self.prec = prec
        self.h = lib().orc_plan_create(N0, N1, len(self.theta), self.theta, m, prec)
        info = np.zeros(7, np.int32)
        sbb = np.zeros(6, np.float32)
        lib().orc_plan_info(self.h, info, sbb)
        self.N0, self.N1, self.nz, self.M, self.n0, self.n1, self.iomit = [int(v) for v in info]
        self.sigma = (float(sbb[0]), float(sbb[3]))
        self.b = (float(sbb[1]), float(sbb[4]))
        self.bf = (float(sbb[2]), float(sbb[5]))

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.orc_plan_destroy(h)
            self.h = None

    def tables(self):
        N0, N1, M = self.N0, self.N1, self.M
        ck = np.zeros(N0 * N1)
        phix = np.zeros(M * TAPS)
        phiy = np.zeros(M * TAPS)
        ix = np.zeros(M * TAPS, np.int32)
        iy = np.zeros(M * TAPS, np.int32)
        w = np.zeros(M, np.float32)
        nodes = np.zeros(2 * M, np.float32)
        ck1 = np.zeros(N0)
        phi1 = np.zeros(N0 * TAPS)
        i1 = np.zeros(N0 * TAPS, np.int32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        lib().orc_plan_tables(self.h, p(ck), p(phix), p(phiy), p(ix), p(iy), p(w), p(nodes), p(ck1), p(phi1), p(i1))
        return dict(ck=ck.reshape(N0, N1), phix=phix.reshape(M, TAPS), phiy=phiy.reshape(M, TAPS),
                    ix=ix.reshape(M, TAPS), iy=iy.reshape(M, TAPS), w=w, nodes=nodes.reshape(M, 2),
                    ck1=ck1, phi1=phi1.reshape(N0, TAPS), i1=i1.reshape(N0, TAPS))

    def forward(self, img):
        f = np.zeros(2 * self.M)
        lib().orc_forward(self.h, _c2d(np.asarray(img).reshape(self.N0, self.N1)), f)
        return _d2c(f, (self.M,))

    def adjoint(self, f):
        out = np.zeros(2 * self.N0 * self.N1)
        lib().orc_adjoint(self.h, _c2d(f), out)
        return _d2c(out, (self.N0, self.N1))

    def sino_to_b(self, sino):
        b = np.zeros(2 * self.M)
        lib().orc_sino_to_b(self.h, np.ascontiguousarray(sino, np.float32).reshape(-1), b)
        return _d2c(b, (self.M,))

    def adjoint1d(self, f):
        out = np.zeros(2 * self.N0)
        lib().orc_adjoint1d(self.h, _c2d(f), out)
        return _d2c(out, (self.N0,))

    def solve(self, sino, pre, icon, post, data_type=2, thr=0.0, m_init=None, trace=False):
        iters = pre + icon + post
        m = np.zeros(self.N0 * self.N1, np.float32)
        res = np.zeros(iters + 1) if trace else None
        al = np.zeros(max(iters, 1)) if trace else None
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
        mi = None if m_init is None else np.ascontiguousarray(m_init, np.float32).reshape(-1)
        lib().orc_icon_solve(self.h, np.ascontiguousarray(sino, np.float32).reshape(-1), pre, icon, post,
                             data_type, thr, m, p(res), p(al), p(mi))
        m = m.reshape(self.N0, self.N1)
        if trace:
            return m, res, al[:iters]
        return m

    def solve_batch(self, sinos, pre, icon, post, data_type=2, thr=0.0, threads=0, trace=False):
        sinos = np.ascontiguousarray(sinos, np.float32)
        S = sinos.shape[0]
        iters = pre + icon + post
        m = np.zeros(S * self.N0 * self.N1, np.float32)
        res = np.zeros(S * (iters + 1)) if trace else None
        used = lib().orc_icon_solve_batch(self.h, S, sinos.reshape(-1), pre, icon, post, data_type, thr, m,
                                          res.ctypes.data_as(ctypes.c_void_p) if trace else None, threads)
        m = m.reshape(S, self.N0, self.N1)
        if trace:
            return m, res.reshape(S, iters + 1), used
        return m, used

    def reproject(self, m):
        out = np.zeros(self.N0, np.float32)
        lib().orc_reproject(self.h, np.ascontiguousarray(m, np.float32).reshape(-1), out)
        return out


def ref_mrc_lib():
    """The reference's own MRC reader/writer compiled from its source (None if not built)."""
    if not os.path.exists(_REF_MRC):
        return None
    return ctypes.CDLL(_REF_MRC)


def preprocess(stack, tilt_deg=None, thickness=0):
    """ICONPreProcess of a stack [nz][ny][nx]: (processed stack, modes[nz])."""
    st = np.ascontiguousarray(stack, np.float32).copy()
    nz, ny, nx = st.shape
    modes = np.zeros(nz)
    tl = None if tilt_deg is None else np.ascontiguousarray(tilt_deg, np.float32)
    lib().orc_preprocess(st.reshape(-1), nx, ny, nz, None if tl is None else tl.ctypes.data_as(ctypes.c_void_p),
                         int(thickness), modes.ctypes.data_as(ctypes.c_void_p))
    return st, modes
