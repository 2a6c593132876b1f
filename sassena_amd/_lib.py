"""ctypes binding of libsassena_amd.so (the C ABI in include/sassena_amd.h).

The library is built in-tree by ``sassena_amd.build``; nothing here falls back
to another implementation: if the shared object is missing or a call fails,
an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsassena_amd.so")

SX_OK = 0
ERRORS = {-1: "SX_ERR_ARG", -2: "SX_ERR_HIP", -3: "SX_ERR_STATE", -4: "SX_ERR_NOMEM",
          -5: "SX_ERR_UNSUPPORTED"}

SCATTERING = {"all": 0, "self": 1}
DSP = {"autocorrelate": 0, "square": 1, "plain": 2}
METHOD = {"fftw": 0, "fft": 0, "direct": 1}
LAYOUT = {"frames": 0, "atoms": 1, "soa": 2}

# Every symbol include/sassena_amd.h declares.
EXPORTS = ("sx_version", "sx_create", "sx_destroy", "sx_last_error", "sx_stage", "sx_set_factors",
           "sx_compute_shell", "sx_compute_shells", "sx_compute_shells_oriented",
           "sx_compute_shells_oriented_noalloc", "sx_compute_multipole",
           "sx_get_timing",
           "sx_pad_length", "sx_batch_create", "sx_batch_destroy", "sx_batch_last_error", "sx_batch_stage",
           "sx_batch_run", "sx_batch_run_one", "sx_batch_get_timing", "sx_batch_time_amplitude",
           "sx_mlp_param_count", "sx_mlp_create",
           "sx_mlp_destroy", "sx_mlp_last_error", "sx_mlp_set_params", "sx_mlp_get_params", "sx_mlp_gradients",
           "sx_mlp_train", "sx_mlp_predict")


class SxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class sx_config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("scattering_type", ctypes.c_int), ("dsp_type", ctypes.c_int),
                ("dsp_method", ctypes.c_int), ("profile", ctypes.c_int), ("vector_block", ctypes.c_int),
                ("atom_split", ctypes.c_int), ("tile_shape", ctypes.c_int),
                ("fft_halves", ctypes.c_int), ("multipole_terms", ctypes.c_int), ("reserved", ctypes.c_int * 6)]


class sx_timing(ctypes.Structure):
    _fields_ = [("amplitude_ms", ctypes.c_double), ("correlate_ms", ctypes.c_double),
                ("finalize_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("amplitude_launches", ctypes.c_size_t), ("correlate_launches", ctypes.c_size_t),
                ("terms", ctypes.c_size_t), ("pad_length", ctypes.c_size_t), ("tile_shape", ctypes.c_int),
                ("shell_progression", ctypes.c_int), ("shell_residuals", ctypes.c_int),
                ("shell_exact_runs", ctypes.c_int)]


_lib = None


def load():
    """Load the library; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m sassena_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, dp, fp = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
    L.sx_version.restype = ctypes.c_char_p
    L.sx_version.argtypes = []
    L.sx_create.argtypes = [ctypes.POINTER(sx_config), ctypes.POINTER(vp)]
    L.sx_destroy.argtypes = [vp]
    L.sx_destroy.restype = None
    L.sx_last_error.argtypes = [vp]
    L.sx_last_error.restype = ctypes.c_char_p
    L.sx_stage.argtypes = [vp, fp, sz, sz, ctypes.c_int]
    L.sx_set_factors.argtypes = [vp, dp, sz]
    L.sx_compute_shell.argtypes = [vp, dp, sz, dp, dp, dp]
    L.sx_compute_shells.argtypes = [vp, sz, dp, sz, dp, dp, dp, dp]
    L.sx_compute_shells_oriented.argtypes = [vp, sz, dp, dp, sz, dp, dp, dp, dp]
    L.sx_compute_shells_oriented_noalloc.argtypes = [vp, sz, dp, dp, sz, dp, dp, dp, dp]
    L.sx_compute_multipole.argtypes = [vp, ctypes.c_int, sz, dp, dp, ctypes.POINTER(ctypes.c_int), sz, dp, dp, dp,
                                       dp]
    L.sx_get_timing.argtypes = [vp, ctypes.POINTER(sx_timing)]
    L.sx_pad_length.argtypes = [sz]
    L.sx_pad_length.restype = sz
    for name in ("sx_create", "sx_stage", "sx_set_factors", "sx_compute_shell", "sx_compute_shells",
                 "sx_compute_shells_oriented", "sx_compute_shells_oriented_noalloc", "sx_compute_multipole",
                 "sx_get_timing"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L
