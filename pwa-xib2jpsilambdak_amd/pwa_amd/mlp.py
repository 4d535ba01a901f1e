"""ctypes face of the BF16-MFMA MLP (include/pwa_mlp.h)."""
import ctypes as C

import numpy as np

from . import abi
from ._lib import PWAError, lib

_fp = C.POINTER(C.c_float)
_bound = False


def _mlp_lib():
    global _bound
    L = lib()
    if not _bound:
        vp = C.c_void_p
        sig = {
            "pwa_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_int,
                                         C.POINTER(vp)]),
            "pwa_mlp_destroy": (C.c_int, [vp]),
            "pwa_mlp_last_error": (C.c_char_p, [vp]),
            "pwa_mlp_n_params": (C.c_int, [vp, C.POINTER(C.c_int64)]),
            "pwa_mlp_get_params": (C.c_int, [vp, _fp]),
            "pwa_mlp_set_params": (C.c_int, [vp, _fp]),
            "pwa_mlp_set_data": (C.c_int, [vp, C.c_int64, _fp, _fp, _fp]),
            "pwa_mlp_loss_grad": (C.c_int, [vp, C.c_int64, C.c_int32, _fp, _fp]),
            "pwa_mlp_train": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]),
            "pwa_mlp_predict": (C.c_int, [vp, C.c_int64, _fp, _fp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return None if a is None else a.ctypes.data_as(_fp)


class MLP:
    """n_in -> [Linear(n_hidden) -> GELU] x n_layers -> Linear(n_out), trained on the GPU."""

    def __init__(self, n_in, n_out, n_hidden=256, n_layers=3, seed=1, device=0):
        self.n_in, self.n_out = n_in, n_out
        self._h = C.c_void_p()
        self._check(_mlp_lib().pwa_mlp_create(n_in, n_hidden, n_layers, n_out, seed, device, C.byref(self._h)),
                    "pwa_mlp_create", ctx=False)
        n = C.c_int64()
        self._check(_mlp_lib().pwa_mlp_n_params(self._h, C.byref(n)), "pwa_mlp_n_params")
        self.n_params = n.value

    def _check(self, rc, what, ctx=True):
        if rc != abi.PWA_OK:
            msg = _mlp_lib().pwa_mlp_last_error(self._h).decode() if ctx and self._h else ""
            raise PWAError(f"{what} failed with status {rc}: {msg}")

    def close(self):
        if self._h:
            _mlp_lib().pwa_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        w = np.zeros(self.n_params, np.float32)
        self._check(_mlp_lib().pwa_mlp_get_params(self._h, _p(w)), "pwa_mlp_get_params")
        return w

    def set_params(self, w):
        w = _f32(w)
        self._check(_mlp_lib().pwa_mlp_set_params(self._h, _p(w)), "pwa_mlp_set_params")

    def set_data(self, x, y, out_weight):
        x, y, ow = _f32(x), _f32(y), _f32(out_weight)
        self._keep = (x, y, ow)
        self._check(_mlp_lib().pwa_mlp_set_data(self._h, x.shape[0], _p(x), _p(y), _p(ow)), "pwa_mlp_set_data")

    def loss_grad(self, first, batch, want_grad=True):
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        self._check(_mlp_lib().pwa_mlp_loss_grad(self._h, first, batch, C.byref(loss), _p(g)), "pwa_mlp_loss_grad")
        return loss.value, g

    def train(self, epochs, batch=256, lr=1e-3, seed=0):
        loss = C.c_float()
        self._check(_mlp_lib().pwa_mlp_train(self._h, epochs, batch, lr, seed, C.byref(loss)), "pwa_mlp_train")
        return loss.value

    def predict(self, x):
        x = _f32(np.atleast_2d(x))
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._check(_mlp_lib().pwa_mlp_predict(self._h, x.shape[0], _p(x), _p(y)), "pwa_mlp_predict")
        return y
