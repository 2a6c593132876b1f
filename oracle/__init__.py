"""FP64 CPU oracle of Sassena's scattering-device hot path (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use
this package, and only as the checker or the timed host path.  The product
library never imports it.

The C restatement lives in ``sx_oracle.c`` (built by ``make -C oracle``); this
module is a thin ctypes front end with numpy arrays.  Each function cites the
reference code it follows in the C source.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsx_oracle.so")
_lib = None

ALL, SELF = 0, 1
DSP = {"autocorrelate": 0, "square": 1, "plain": 2}
METHOD = {"fftw": 0, "fft": 0, "direct": 1}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile libsx_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sz = ctypes.c_size_t
        L.sxo_fft.argtypes = [_dp, sz, ctypes.c_int]
        L.sxo_autocorrelate_fft.argtypes = [_dp, sz]
        L.sxo_autocorrelate_direct.argtypes = [_dp, sz]
        L.sxo_square_elements.argtypes = [_dp, sz]
        L.sxo_store_one.argtypes = [_dp, sz, _dp]
        L.sxo_scatter_all.argtypes = [_fp, sz, sz, _dp, _dp, _dp]
        L.sxo_scatter_self.argtypes = [_fp, sz, sz, ctypes.c_double, _dp, _dp]
        for fn in (L.sxo_compute_all, L.sxo_compute_self):
            fn.argtypes = [_fp, sz, sz, _dp, _dp, sz, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, _dp, _dp, _dp]
        L.sxo_init_subvectors.argtypes = [_dp, _dp, sz, _dp]
        L.sxo_init_subvectors.restype = sz
        L.sxo_normalise_rows.argtypes = [_dp, sz]
        L.sxo_scan.argtypes = [ctypes.c_double, ctypes.c_double, sz,
                               ctypes.c_double, _dp, _dp]
        L.sxo_scan.restype = sz
        ip = ctypes.POINTER(ctypes.c_int)
        L.sxo_sph_bessel_all.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.sxo_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.sxo_cyl_bessel_j.argtypes = [ctypes.c_int, ctypes.c_double]
        L.sxo_cyl_bessel_j.restype = ctypes.c_double
        L.sxo_scatter_mpsphere.argtypes = [_fp, sz, sz, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
        L.sxo_scatter_mpcylinder.argtypes = [_fp, sz, sz, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.sxo_cylinder_base.argtypes = [_dp, _dp]
        L.sxo_compute_multipole.argtypes = [ctypes.c_int, _fp, sz, sz, _dp, _dp, _dp, ip, sz, ctypes.c_int,
                                            ctypes.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _cplx_to_interleaved(x):
    x = np.asarray(x, dtype=np.complex128)
    out = np.empty(2 * x.size, dtype=np.float64)
    out[0::2] = x.real.ravel()
    out[1::2] = x.imag.ravel()
    return out


def _interleaved_to_cplx(a):
    return a[0::2] + 1j * a[1::2]


def fft(x, sign=-1):
    """Unnormalised DFT (sign -1: FFTW_FORWARD, +1: FFTW_BACKWARD)."""
    buf = _cplx_to_interleaved(x)
    if lib().sxo_fft(_d(buf), len(buf) // 2, sign):
        raise RuntimeError("sxo_fft failed")
    return _interleaved_to_cplx(buf)


def autocorrelate_fft(a):
    """smath.cpp:141-156 on a length-NF signal (padded here to 2NF)."""
    a = np.asarray(a, dtype=np.complex128)
    NF = a.size
    buf = _cplx_to_interleaved(np.concatenate([a, np.zeros(NF, np.complex128)]))
    if lib().sxo_autocorrelate_fft(_d(buf), NF):
        raise RuntimeError("sxo_autocorrelate_fft failed")
    return _interleaved_to_cplx(buf)[:NF]


def autocorrelate_direct(a):
    """smath.cpp:51-76 (the conjugate of autocorrelate_fft's result)."""
    buf = _cplx_to_interleaved(a)
    if lib().sxo_autocorrelate_direct(_d(buf), len(buf) // 2):
        raise RuntimeError("sxo_autocorrelate_direct failed")
    return _interleaved_to_cplx(buf)


def scatter_all(coords_frames, b, q):
    """A(t) for one vector, coords frames-major [NF][NA][3] float32."""
    c = np.ascontiguousarray(coords_frames, dtype=np.float32)
    NF, NA = c.shape[0], c.shape[1]
    bb = np.ascontiguousarray(b, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros(2 * NF, np.float64)
    lib().sxo_scatter_all(_f(c), NA, NF, _d(bb), _d(qq), _d(out))
    return _interleaved_to_cplx(out)


def compute(kind, coords, b, qv, dsp="autocorrelate", method="fftw", nthreads=1):
    """One q shell through the reference device.

    kind: "all" (coords frames-major [NF][NA][3]) or "self" (atom-major
    [NA][NF][3]).  qv: [NM][3] orientation-scaled q vectors.
    Returns (fqt complex[NF], fq complex, fq2 complex).
    """
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if kind == "all":
        NF, NA = c.shape[0], c.shape[1]
        fn = lib().sxo_compute_all
    elif kind == "self":
        NA, NF = c.shape[0], c.shape[1]
        fn = lib().sxo_compute_self
    else:
        raise ValueError(kind)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    q = np.ascontiguousarray(np.asarray(qv, dtype=np.float64).reshape(-1, 3))
    fqt = np.zeros(2 * NF, np.float64)
    fq = np.zeros(2, np.float64)
    fq2 = np.zeros(2, np.float64)
    err = fn(_f(c), NA, NF, _d(bb), _d(q), q.shape[0], DSP[dsp], METHOD[method],
             int(nthreads), _d(fqt), _d(fq), _d(fq2))
    if err:
        raise RuntimeError(f"oracle compute failed ({err})")
    return _interleaved_to_cplx(fqt), complex(fq[0], fq[1]), complex(fq2[0], fq2[1])


def init_subvectors(q, u=None):
    """abstract_vectors_scatter_device.cpp:112-175 (none/file/sphere)."""
    qq = np.ascontiguousarray(q, dtype=np.float64)
    if u is None or len(u) == 0:
        uu = np.zeros(3, np.float64)
        n = 0
    else:
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        n = uu.shape[0]
    out = np.zeros((max(n, 1), 3), np.float64)
    k = lib().sxo_init_subvectors(_d(qq), _d(uu), n, _d(out))
    return out[:k]


def normalise_rows(v):
    """parameters.cpp:932-944."""
    vv = np.array(v, dtype=np.float64).reshape(-1, 3)
    lib().sxo_normalise_rows(_d(vv), vv.shape[0])
    return vv


def scan(frm, to, points, exponent=1.0, base=(1.0, 0.0, 0.0)):
    """parameters.cpp:1142-1162 (single scan)."""
    bb = np.ascontiguousarray(base, dtype=np.float64)
    out = np.zeros((max(points, 1), 3), np.float64)
    k = lib().sxo_scan(frm, to, points, exponent, _d(bb), _d(out))
    return out[:k]


def sph_bessel_all(L, x):
    out = np.zeros(L + 1)
    lib().sxo_sph_bessel_all(L, x, _d(out))
    return out


def ylm(l, m, theta, phi):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().sxo_ylm(l, m, theta, phi, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def cyl_bessel_j(n, x):
    return lib().sxo_cyl_bessel_j(n, x)


def sphere_moments(L):
    """parameters.cpp:1058-1066: (0,0), then l = 1..L with m = -l..l."""
    return [(0, 0)] + [(l, m) for l in range(1, L + 1) for m in range(-l, l + 1)]


def cylinder_moments(L):
    """parameters.cpp:1068-1075: (0,0), then l = 1..L with m = 0..3."""
    return [(0, 0)] + [(l, m) for l in range(1, L + 1) for m in range(4)]


def scatter_multipole(kind, coords_frames, b, q, l, m, axis=(0.0, 0.0, 1.0)):
    """A_lm(t) of one moment (multipole_scatter_device.cpp:428-498 sphere,
    900-987 cylinder); coords frames-major [NF][NA][3]; q: the q vector
    (the sphere uses its length only)."""
    c = np.ascontiguousarray(coords_frames, dtype=np.float32)
    NF, NA = c.shape[0], c.shape[1]
    bb = np.ascontiguousarray(b, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros(2 * NF)
    if kind == "sphere":
        lib().sxo_scatter_mpsphere(_f(c), NA, NF, _d(bb), float(np.linalg.norm(qq)), int(l), int(m), _d(out))
    else:
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        lib().sxo_scatter_mpcylinder(_f(c), NA, NF, _d(bb), _d(qq), _d(ax), int(l), int(m), _d(out))
    return _interleaved_to_cplx(out)


def cylinder_base(axis):
    """CartesianVectorBase(axis) (coor3d.cpp:278-296): rows e_r, e_phi, e_z."""
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    out = np.zeros(9)
    lib().sxo_cylinder_base(_d(ax), _d(out))
    return out.reshape(3, 3)


def compute_multipole(kind, coords_frames, b, q, moments, axis=(0.0, 0.0, 1.0), dsp="autocorrelate",
                      method="fftw"):
    """One shell of MPSphereScatterDevice (kind "sphere") or
    MPCylinderScatterDevice ("cylinder"); coords frames-major [NF][NA][3]."""
    c = np.ascontiguousarray(coords_frames, dtype=np.float32)
    NF, NA = c.shape[0], c.shape[1]
    bb = np.ascontiguousarray(b, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    mo = np.ascontiguousarray(np.asarray(moments, dtype=np.int32).reshape(-1, 2))
    fqt = np.zeros(2 * NF)
    fq = np.zeros(2)
    fq2 = np.zeros(2)
    err = lib().sxo_compute_multipole({"sphere": 0, "cylinder": 1}[kind], _f(c), NA, NF, _d(bb), _d(qq), _d(ax),
                                      mo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), mo.shape[0], DSP[dsp],
                                      METHOD[method], _d(fqt), _d(fq), _d(fq2))
    if err:
        raise RuntimeError(f"oracle multipole failed ({err})")
    return _interleaved_to_cplx(fqt), complex(fq[0], fq[1]), complex(fq2[0], fq2[1])
