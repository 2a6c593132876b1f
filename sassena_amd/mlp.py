"""The control-variate network on the GPU (sx_mlp_* of the C ABI)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SxError

_fp = ctypes.POINTER(ctypes.c_float)
IN = 16


def _bind(L):
    if getattr(L, "_sx_mlp_bound", False):
        return L
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.sx_mlp_param_count.argtypes = []
    L.sx_mlp_param_count.restype = sz
    L.sx_mlp_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.sx_mlp_destroy.argtypes = [vp]
    L.sx_mlp_destroy.restype = None
    L.sx_mlp_last_error.argtypes = [vp]
    L.sx_mlp_last_error.restype = ctypes.c_char_p
    L.sx_mlp_set_params.argtypes = [vp, _fp, sz]
    L.sx_mlp_get_params.argtypes = [vp, _fp, sz]
    L.sx_mlp_gradients.argtypes = [vp, _fp, _fp, sz, _fp, _fp]
    L.sx_mlp_train.argtypes = [vp, _fp, _fp, sz, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong, _fp]
    L.sx_mlp_predict.argtypes = [vp, _fp, sz, _fp]
    for n in ("sx_mlp_create", "sx_mlp_set_params", "sx_mlp_get_params", "sx_mlp_gradients", "sx_mlp_train",
              "sx_mlp_predict"):
        getattr(L, n).restype = ctypes.c_int
    L._sx_mlp_bound = True
    return L


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class MLP:
    def __init__(self, device=0, seed=0):
        self._L = _bind(_lib.load())
        h = ctypes.c_void_p()
        rc = self._L.sx_mlp_create(device, seed, ctypes.byref(h))
        if rc != _lib.SX_OK:
            raise SxError(rc, "sx_mlp_create")
        self._h = h
        self.nparam = int(self._L.sx_mlp_param_count())

    def _check(self, rc, what):
        if rc != _lib.SX_OK:
            raise SxError(rc, f"{what}: {self._L.sx_mlp_last_error(self._h).decode()}")

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        self._check(self._L.sx_mlp_get_params(self._h, p.ctypes.data_as(_fp), p.size), "sx_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p, (self.nparam,))
        self._check(self._L.sx_mlp_set_params(self._h, p.ctypes.data_as(_fp), p.size), "sx_mlp_set_params")

    def gradients(self, x, t):
        x = _f32(x)
        t = _f32(t)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_float()
        self._check(self._L.sx_mlp_gradients(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), t.size,
                                             g.ctypes.data_as(_fp), ctypes.byref(loss)), "sx_mlp_gradients")
        return g, loss.value

    def train(self, x, t, steps=2000, batch=4096, lr=2e-3, seed=1):
        x = _f32(x)
        t = _f32(t)
        loss = ctypes.c_float()
        self._check(self._L.sx_mlp_train(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), t.size, steps, batch,
                                         lr, seed, ctypes.byref(loss)), "sx_mlp_train")
        return loss.value

    def predict(self, x):
        x = _f32(x)
        y = np.zeros(x.shape[0], np.float32)
        self._check(self._L.sx_mlp_predict(self._h, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_fp)),
                    "sx_mlp_predict")
        return y

    def close(self):
        if getattr(self, "_h", None):
            self._L.sx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
