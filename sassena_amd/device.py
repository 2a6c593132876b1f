"""Host-side mirror of Sassena's scattering-device interface over the C ABI.

``ScatterDevice`` plays the role of ``IScatterDevice`` created by
``ScatterDeviceFactory::create`` (scatter_device_factory.cpp:145-200): it is
configured by scattering.type / dsp.type / dsp.method, stages a coordinate
block (DataStagerByFrame / DataStagerByAtom layouts, data_stager.cpp:72-118,
214-338), takes the scattering factors (ScatterFactors::update,
scatter_factors.cpp:56-78) and computes one shell per ``compute_shell`` call,
returning (fqt, fq, fq2) as AbstractScatterDevice::write hands them to the
signal writer (abstract_scatter_device.cpp:239-244; fq0 = fqt[0]).
All arithmetic runs in libsassena_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SxError


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cx(a):
    """(..., 2) float64 (re, im) pairs as complex128 values, without a copy."""
    return a.view(np.complex128)[..., 0]


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class ScatterDevice:
    """One staged trajectory block on one GPU (not thread-safe)."""

    def __init__(self, scattering="all", dsp="autocorrelate", method="fftw", device=0, profile=False,
                 vector_block=0, atom_split=0, tile_shape=0, fft_halves=0, multipole_terms=0):
        L = _lib.load()
        self._L = L
        cfg = _lib.sx_config()
        cfg.device = device
        cfg.scattering_type = _lib.SCATTERING[scattering]
        cfg.dsp_type = _lib.DSP[dsp]
        cfg.dsp_method = _lib.METHOD[method]
        cfg.profile = 1 if profile else 0
        cfg.vector_block = vector_block
        cfg.atom_split = atom_split
        cfg.tile_shape = tile_shape
        cfg.fft_halves = fft_halves
        cfg.multipole_terms = multipole_terms
        h = ctypes.c_void_p()
        rc = L.sx_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != _lib.SX_OK:
            raise SxError(rc, f"sx_create(scattering={scattering}, dsp={dsp}, method={method})")
        self._h = h
        self.scattering, self.dsp, self.method = scattering, dsp, method
        self.NA = self.NF = 0

    def _check(self, rc, what):
        if rc != _lib.SX_OK:
            msg = self._L.sx_last_error(self._h).decode()
            raise SxError(rc, f"{what}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.sx_destroy(self._h)
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

    # ------------------------------------------------------------ staging
    def stage(self, coords, layout="soa"):
        """layout 'soa' [3][NA][NF], 'frames' [NF][NA][3] or 'atoms' [NA][NF][3] (float32)."""
        c = np.ascontiguousarray(coords, dtype=np.float32)
        if c.ndim != 3:
            raise ValueError("coordinates must be a 3-d array")
        if layout == "soa":
            NA, NF = c.shape[1], c.shape[2]
        elif layout == "frames":
            NF, NA = c.shape[0], c.shape[1]
        elif layout == "atoms":
            NA, NF = c.shape[0], c.shape[1]
        else:
            raise ValueError(layout)
        self._check(self._L.sx_stage(self._h, _fp(c), NA, NF, _lib.LAYOUT[layout]), "sx_stage")
        self.NA, self.NF = NA, NF

    def set_factors(self, b):
        bb = np.ascontiguousarray(b, dtype=np.float64)
        self._check(self._L.sx_set_factors(self._h, _dp(bb), bb.size), "sx_set_factors")

    # ------------------------------------------------------------ compute
    def compute_shell(self, qv):
        """qv [NM][3] -> (fqt complex[NF], fq complex, fq2 complex)."""
        fqt, fq, fq2 = self.compute_shells(np.asarray(qv, dtype=np.float64)[None])
        return fqt[0], fq[0], fq2[0]

    def compute_shells(self, qv, b_shell=None):
        """qv [nshells][NM][3] -> (fqt [nshells][NF], fq [nshells], fq2 [nshells]), complex."""
        q = np.ascontiguousarray(qv, dtype=np.float64)
        if q.ndim != 3 or q.shape[2] != 3:
            raise ValueError("qv must be [nshells][NM][3]")
        ns, NM = q.shape[0], q.shape[1]
        fqt = np.zeros((ns, self.NF, 2), np.float64)
        fq = np.zeros((ns, 2), np.float64)
        fq2 = np.zeros((ns, 2), np.float64)
        bp = None
        if b_shell is not None:
            bs = np.ascontiguousarray(b_shell, dtype=np.float64)
            if bs.shape != (ns, self.NA):
                raise ValueError("b_shell must be [nshells][NA]")
            bp = _dp(bs)
        self._check(self._L.sx_compute_shells(self._h, ns, _dp(q), NM, bp, _dp(fqt), _dp(fq), _dp(fq2)),
                    "sx_compute_shells")
        return (_cx(fqt), _cx(fq), _cx(fq2))

    def compute_oriented(self, qlen, u, b_shell=None, noalloc=False):
        """Shells sharing one orientation set: q_{s,m} = qlen[s] * u[m].

        qlen [nshells], u [NM][3] unit rows -> (fqt [nshells][NF], fq, fq2).
        noalloc: through sx_compute_shells_oriented_noalloc (work memory in
        the library's fixed per-GPU scratch, no device allocation)."""
        ql = np.ascontiguousarray(qlen, dtype=np.float64).ravel()
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        ns, NM = ql.size, uu.shape[0]
        fqt = np.zeros((ns, self.NF, 2), np.float64)
        fq = np.zeros((ns, 2), np.float64)
        fq2 = np.zeros((ns, 2), np.float64)
        bp = None
        if b_shell is not None:
            bs = np.ascontiguousarray(b_shell, dtype=np.float64)
            if bs.shape != (ns, self.NA):
                raise ValueError("b_shell must be [nshells][NA]")
            bp = _dp(bs)
        fn = self._L.sx_compute_shells_oriented_noalloc if noalloc else self._L.sx_compute_shells_oriented
        self._check(fn(self._h, ns, _dp(ql), _dp(uu), NM, bp, _dp(fqt), _dp(fq), _dp(fq2)),
                    "sx_compute_shells_oriented" + ("_noalloc" if noalloc else ""))
        return (_cx(fqt), _cx(fq), _cx(fq2))

    def compute_multipole(self, kind, qv, moments, axis=(0.0, 0.0, 1.0), b_shell=None):
        """Multipole averaging: kind "sphere" or "cylinder"; qv [nshells][3];
        moments [[l, m], ...] -> (fqt [nshells][NF], fq, fq2)."""
        q = np.ascontiguousarray(qv, dtype=np.float64).reshape(-1, 3)
        mo = np.ascontiguousarray(np.asarray(moments, dtype=np.int32).reshape(-1, 2))
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        ns = q.shape[0]
        fqt = np.zeros((ns, self.NF, 2), np.float64)
        fq = np.zeros((ns, 2), np.float64)
        fq2 = np.zeros((ns, 2), np.float64)
        bp = None
        if b_shell is not None:
            bs = np.ascontiguousarray(b_shell, dtype=np.float64)
            if bs.shape != (ns, self.NA):
                raise ValueError("b_shell must be [nshells][NA]")
            bp = _dp(bs)
        kinds = {"sphere": 0, "cylinder": 1}
        self._check(self._L.sx_compute_multipole(self._h, kinds[kind], ns, _dp(q), _dp(ax),
                                                 mo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), mo.shape[0], bp,
                                                 _dp(fqt), _dp(fq), _dp(fq2)), "sx_compute_multipole")
        return (_cx(fqt), _cx(fq), _cx(fq2))

    def timing(self):
        t = _lib.sx_timing()
        self._check(self._L.sx_get_timing(self._h, ctypes.byref(t)), "sx_get_timing")
        return {k: getattr(t, k) for k, _ in t._fields_}


def pad_length(NF):
    return int(_lib.load().sx_pad_length(NF))


def version():
    return _lib.load().sx_version().decode()
