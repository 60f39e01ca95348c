"""ctypes mirror of the warm-start U-Net C ABI (include/icon_unet.h)."""
import ctypes

import numpy as np

from .icon import DeviceArray, lib


class UNetError(RuntimeError):
    pass


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.icon_unet_last_error.restype = ctypes.c_char_p
        sig = {
            "icon_unet_create": [ctypes.POINTER(vp), ci, ci, ci, ci, ci],
            "icon_unet_destroy": [vp],
            "icon_unet_param_count": [vp, ctypes.POINTER(ctypes.c_int64)],
            "icon_unet_init": [vp, ctypes.c_uint64],
            "icon_unet_get_params": [vp, vp, ctypes.c_int64],
            "icon_unet_set_params": [vp, vp, ctypes.c_int64],
            "icon_unet_forward": [vp, ci, vp, vp, vp],
            "icon_unet_warm_start": [vp, ci, vp, vp, vp],
            "icon_unet_set_output_gain": [vp, ctypes.c_float],
            "icon_unet_gradients": [vp, ci, vp, vp, ctypes.POINTER(ctypes.c_double), vp],
            "icon_unet_get_gradients": [vp, vp, ctypes.c_int64],
            "icon_unet_train_step": [vp, ci, vp, vp, ctypes.c_float, ctypes.POINTER(ctypes.c_double), vp],
        }
        for k, a in sig.items():
            getattr(L, k).argtypes = a
            getattr(L, k).restype = ci
        _bound = True
    return L


def _check(rc, what):
    if rc:
        raise UNetError("%s: %s" % (what, _lib().icon_unet_last_error().decode(errors="replace")))


def _p(x):
    return x.ptr if isinstance(x, DeviceArray) else ctypes.c_void_p(int(x))


def _s(stream):
    """A stream argument: None (the default stream), a Stream or a raw handle."""
    if stream is None:
        return None
    h = getattr(stream, "h", None)
    return h if h is not None else ctypes.c_void_p(int(stream))


class UNet:
    """U-Net on one device for batches of up to max_batch images of H x W."""

    def __init__(self, H, W, max_batch=1, train=False, device=0, seed=None):
        L = _lib()
        self.h = ctypes.c_void_p()
        _check(L.icon_unet_create(ctypes.byref(self.h), device, max_batch, H, W, 1 if train else 0), "create")
        self.H, self.W, self.max_batch, self.train = H, W, max_batch, train
        n = ctypes.c_int64()
        _check(L.icon_unet_param_count(self.h, ctypes.byref(n)), "param_count")
        self.nparam = n.value
        if seed is not None:
            _check(L.icon_unet_init(self.h, seed), "init")

    def close(self):
        if self.h and self.h.value:
            _lib().icon_unet_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        _check(_lib().icon_unet_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p), self.nparam), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(_lib().icon_unet_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p), self.nparam), "set_params")

    def forward_dev(self, B, d_in, d_out, stream=None):
        _check(_lib().icon_unet_forward(self.h, B, _p(d_in), _p(d_out), _s(stream)), "forward")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        B = x.shape[0]
        d_in = DeviceArray.from_host(x)
        d_out = DeviceArray(x.shape, np.float32)
        self.forward_dev(B, d_in, d_out)
        return d_out.numpy()

    def set_output_gain(self, gain):
        """Factor g on the warm start's result: g * (m + s * U(m / s)) (default 1)."""
        _check(_lib().icon_unet_set_output_gain(self.h, float(gain)), "set_output_gain")

    def warm_start_dev(self, S, d_m, d_out, stream=None):
        """d_out = m + s * U(m / s) per slice, s = RMS of the slice (device buffers),
        times the output gain."""
        _check(_lib().icon_unet_warm_start(self.h, S, _p(d_m), _p(d_out), _s(stream)), "warm_start")

    def warm_start(self, m):
        m = np.ascontiguousarray(m, np.float32)
        d_m = DeviceArray.from_host(m)
        d_o = DeviceArray(m.shape, np.float32)
        self.warm_start_dev(m.shape[0], d_m, d_o)
        return d_o.numpy()

    def gradients(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        d_in = DeviceArray.from_host(x)
        d_t = DeviceArray.from_host(np.ascontiguousarray(target, np.float32))
        loss = ctypes.c_double()
        _check(_lib().icon_unet_gradients(self.h, x.shape[0], d_in.ptr, d_t.ptr, ctypes.byref(loss), None), "gradients")
        g = np.zeros(self.nparam, np.float32)
        _check(_lib().icon_unet_get_gradients(self.h, g.ctypes.data_as(ctypes.c_void_p), self.nparam), "get_gradients")
        return loss.value, g

    def train_step_dev(self, B, d_in, d_target, lr, stream=None):
        loss = ctypes.c_double()
        _check(_lib().icon_unet_train_step(self.h, B, _p(d_in), _p(d_target), lr, ctypes.byref(loss), stream),
               "train_step")
        return loss.value
