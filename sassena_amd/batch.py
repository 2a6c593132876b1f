"""Batched trajectories over the C ABI (sx_batch_*): many uneven coherent
trajectories through one captured HIP graph with a persistent amplitude
kernel.  Per trajectory the results are those of
``ScatterDevice.compute_oriented``."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SxError

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class sx_trajectory(ctypes.Structure):
    _fields_ = [("xyz", _fp), ("NA", ctypes.c_size_t), ("NF", ctypes.c_size_t), ("layout", ctypes.c_int),
                ("b", _dp), ("qlen", _dp), ("nshells", ctypes.c_size_t), ("u", _dp), ("NM", ctypes.c_size_t),
                ("fqt", _dp), ("fq", _dp), ("fq2", _dp)]


class sx_batch_timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("amplitude_ms", ctypes.c_double), ("terms", ctypes.c_size_t),
                ("tiles", ctypes.c_size_t), ("tile_shape", ctypes.c_int), ("terms_stepped", ctypes.c_size_t),
                ("terms_residual", ctypes.c_size_t), ("terms_exact_runs", ctypes.c_size_t)]


def _bind(L):
    if getattr(L, "_sx_batch_bound", False):
        return L
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.sx_batch_create.argtypes = [ctypes.POINTER(_lib.sx_config), ctypes.POINTER(vp)]
    L.sx_batch_create.restype = ctypes.c_int
    L.sx_batch_destroy.argtypes = [vp]
    L.sx_batch_destroy.restype = None
    L.sx_batch_last_error.argtypes = [vp]
    L.sx_batch_last_error.restype = ctypes.c_char_p
    L.sx_batch_stage.argtypes = [vp, ctypes.POINTER(sx_trajectory), sz]
    L.sx_batch_stage.restype = ctypes.c_int
    L.sx_batch_run.argtypes = [vp]
    L.sx_batch_run.restype = ctypes.c_int
    L.sx_batch_run_one.argtypes = [vp, sz, ctypes.POINTER(ctypes.c_double)]
    L.sx_batch_run_one.restype = ctypes.c_int
    L.sx_batch_get_timing.argtypes = [vp, ctypes.POINTER(sx_batch_timing)]
    L.sx_batch_get_timing.restype = ctypes.c_int
    L.sx_batch_time_amplitude.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.sx_batch_time_amplitude.restype = ctypes.c_int
    L._sx_batch_bound = True
    return L


class Batch:
    """Trajectories given as (xyz [3][NA][NF] float32, b [NA], qlen [ns], u [NM][3])."""

    def __init__(self, device=0, method="fftw", tile_shape=0):
        self._L = _bind(_lib.load())
        cfg = _lib.sx_config()
        cfg.device = device
        cfg.dsp_method = _lib.METHOD[method]
        cfg.tile_shape = tile_shape
        h = ctypes.c_void_p()
        rc = self._L.sx_batch_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != _lib.SX_OK:
            raise SxError(rc, "sx_batch_create")
        self._h = h
        self._keep = []
        self.outputs = []

    def _check(self, rc, what):
        if rc != _lib.SX_OK:
            raise SxError(rc, f"{what}: {self._L.sx_batch_last_error(self._h).decode()}")

    def stage(self, trajectories):
        arr = (sx_trajectory * len(trajectories))()
        self._keep = []
        self.outputs = []
        for k, (xyz, b, qlen, u) in enumerate(trajectories):
            x = np.ascontiguousarray(xyz, dtype=np.float32)
            bb = np.ascontiguousarray(b, dtype=np.float64)
            ql = np.ascontiguousarray(qlen, dtype=np.float64).ravel()
            uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
            NA, NF = x.shape[1], x.shape[2]
            fqt = np.zeros((ql.size, NF, 2))
            fq = np.zeros((ql.size, 2))
            fq2 = np.zeros((ql.size, 2))
            self._keep += [x, bb, ql, uu]
            self.outputs.append((fqt, fq, fq2))
            t = arr[k]
            t.xyz = x.ctypes.data_as(_fp)
            t.NA, t.NF, t.layout = NA, NF, _lib.LAYOUT["soa"]
            t.b = bb.ctypes.data_as(_dp)
            t.qlen, t.nshells = ql.ctypes.data_as(_dp), ql.size
            t.u, t.NM = uu.ctypes.data_as(_dp), uu.shape[0]
            t.fqt, t.fq, t.fq2 = fqt.ctypes.data_as(_dp), fq.ctypes.data_as(_dp), fq2.ctypes.data_as(_dp)
        self._arr = arr
        self._check(self._L.sx_batch_stage(self._h, arr, len(trajectories)), "sx_batch_stage")

    def _results(self, k):
        fqt, fq, fq2 = self.outputs[k]
        # copies: the (re, im) buffers are the next run's outputs
        return tuple(a.view(np.complex128)[..., 0].copy() for a in (fqt, fq, fq2))

    def run(self):
        self._check(self._L.sx_batch_run(self._h), "sx_batch_run")
        return [self._results(k) for k in range(len(self.outputs))]

    def run_one(self, k):
        ms = ctypes.c_double()
        self._check(self._L.sx_batch_run_one(self._h, k, ctypes.byref(ms)), "sx_batch_run_one")
        return self._results(k), ms.value

    def time_amplitude(self):
        """Device time (ms) of the batch's amplitude kernel alone."""
        ms = ctypes.c_double()
        self._check(self._L.sx_batch_time_amplitude(self._h, ctypes.byref(ms)), "sx_batch_time_amplitude")
        return ms.value

    def timing(self):
        t = sx_batch_timing()
        self._check(self._L.sx_batch_get_timing(self._h, ctypes.byref(t)), "sx_batch_get_timing")
        return {k: getattr(t, k) for k, _ in t._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._L.sx_batch_destroy(self._h)
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
