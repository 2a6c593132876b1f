"""Learned control variate for the orientational average.

The plain estimate of a coherent shell average uses NM orientations (1,000 in
the large case); its cost is linear in NM.  Here a small network, trained on
the GPU by libsassena_amd.so's own MFMA kernels, maps cheap features of a
shell and lag -- |q|, log lag, the 16-orientation estimate F16(q, tau) and
its neighbouring shells -- to the many-orientation average.  It is trained on
the first part of the trajectory, where both estimates are computed.

On the remaining frames the network output g is used as a control variate of
the 16-orientation estimate,

    F_cv(q, tau) = F16(q, tau) + c_q (g(q, tau) - F16(q, tau)),

with the coefficient c_q in [0, 1] fitted per shell on held-out training
frames (the network is trained on the first half of the training frames and
c_q is the least-squares coefficient on the second half).  c_q = 0 returns
the plain 16-orientation estimate, so the network only acts where it is
shown to reduce the error on data it was not trained on.  The plain
many-orientation average stays the fallback and is what the result is judged
against (``evaluate``).

All values are normalised by nb = sum_j b_j^2 (the large-|q| limit of
F(q, 0) for uncorrelated positions), which is known exactly from the
scattering factors.  Lags beyond the training range are presented to the
network at the last trained lag.
"""
from __future__ import annotations

import time

import numpy as np

from .device import ScatterDevice
from .mlp import MLP

NFEAT = 16


def features(qlen, f16, nb, tmax, qmax):
    """Rows (shell, lag) of the network input, zero padded to 16 columns:
    |q|/qmax, log1p(tc)/log1p(tmax), tc/tmax (tc = lag clamped to tmax - 1),
    Re F16/nb, Im F16/nb, Re F16(q,0)/nb, the mean Re F16/nb of the two
    neighbouring shells, (|q|/qmax)^2 log1p(tc)/log1p(tmax)."""
    ns, NF = f16.shape
    tc = np.minimum(np.arange(NF, dtype=np.float64), tmax - 1.0)
    q = np.repeat(np.asarray(qlen, np.float64) / qmax, NF)
    lt = np.tile(np.log1p(tc) / np.log1p(tmax), ns)
    re = f16.real / nb
    up = re[np.minimum(np.arange(ns) + 1, ns - 1)]
    dn = re[np.maximum(np.arange(ns) - 1, 0)]
    x = np.zeros((ns * NF, NFEAT), np.float32)
    x[:, 0] = q
    x[:, 1] = lt
    x[:, 2] = np.tile(tc / tmax, ns)
    x[:, 3] = re.ravel()
    x[:, 4] = (f16.imag / nb).ravel()
    x[:, 5] = np.repeat(re[:, 0], NF)
    x[:, 6] = (0.5 * (up + dn)).ravel()
    x[:, 7] = q * q * lt
    return x


def shell_errors(f, ref):
    """Per shell: max and mean over lags of |f - ref| / |ref(0)|."""
    d = np.abs(f - ref) / np.abs(ref[:, :1])
    return d.max(axis=1), d.mean(axis=1)


def cv_coefficients(g, f16, fmany):
    """Least-squares c_q of fmany - f16 ~ c_q (g - f16) per shell (real
    parts), clipped to [0, 1]."""
    d = g - f16.real
    num = np.sum(d * (fmany.real - f16.real), axis=1)
    den = np.sum(d * d, axis=1)
    c = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
    return np.clip(c, 0.0, 1.0)


def run(xyz, b, qlen, u, ntrain_frames, n_small=16, steps=3000, batch=4096, lr=2e-3, seed=1, device=0):
    """The full pipeline on one trajectory (xyz [3][NA][NF] float32).

    Returns a dict with the estimates on the evaluation frames (the frames
    after ntrain_frames), the many-orientation reference there, timings
    (seconds) of each stage, the per-shell coefficients and the errors of
    both estimates."""
    qlen = np.asarray(qlen, np.float64)
    u = np.asarray(u, np.float64)
    nb = float(np.sum(np.asarray(b, np.float64) ** 2))
    NF = xyz.shape[2]
    qmax = float(np.max(qlen))
    half = ntrain_frames // 2
    seg_a = np.ascontiguousarray(xyz[:, :, :half])
    seg_b = np.ascontiguousarray(xyz[:, :, half:ntrain_frames])
    evalf = np.ascontiguousarray(xyz[:, :, ntrain_frames:])
    tmax = float(half)
    usmall = u[:n_small]
    tm = {}

    def shells(seg, orient):
        with ScatterDevice("all", device=device) as dev:
            dev.stage(seg, layout="soa")
            dev.set_factors(b)
            t0 = time.perf_counter()
            fqt, _, _ = dev.compute_oriented(qlen, orient)
            return fqt, time.perf_counter() - t0

    # training data: both estimates on the first half of the training frames
    f_many_a, t1 = shells(seg_a, u)
    f_small_a, t2 = shells(seg_a, usmall)
    x_a = features(qlen, f_small_a, nb, tmax, qmax)
    y_a = (f_many_a.real / nb).ravel().astype(np.float32)
    net = MLP(device=device, seed=seed)
    t0 = time.perf_counter()
    loss = net.train(x_a, y_a, steps=steps, batch=batch, lr=lr, seed=seed)
    tm["train"] = time.perf_counter() - t0
    # control-variate coefficients on the held-out second half
    f_many_b, t3 = shells(seg_b, u)
    f_small_b, t4 = shells(seg_b, usmall)
    g_b = nb * net.predict(features(qlen, f_small_b, nb, tmax, qmax)).reshape(f_small_b.shape).astype(np.float64)
    coef = cv_coefficients(g_b, f_small_b, f_many_b)
    tm["many_train"] = t1 + t3
    tm["small_train"] = t2 + t4

    # evaluation frames: the cheap estimate and its correction
    f_small_ev, tm["small_eval"] = shells(evalf, usmall)
    t0 = time.perf_counter()
    g_ev = nb * net.predict(features(qlen, f_small_ev, nb, tmax, qmax)).reshape(f_small_ev.shape).astype(np.float64)
    f_cv = f_small_ev + coef[:, None] * (g_ev - f_small_ev.real)
    tm["predict"] = time.perf_counter() - t0
    # the judge: the plain many-orientation average on the same frames
    f_many_ev, tm["many_eval"] = shells(evalf, u)
    net.close()

    e_small_max, e_small_mean = shell_errors(f_small_ev, f_many_ev)
    e_cv_max, e_cv_mean = shell_errors(f_cv, f_many_ev)
    e_g_max, e_g_mean = shell_errors(g_ev + 1j * f_small_ev.imag, f_many_ev)
    bias = float(np.mean((f_cv.real - f_many_ev.real) / np.abs(f_many_ev[:, :1].real)))
    return {
        "f_cv": f_cv, "f_small": f_small_ev, "f_many": f_many_ev, "g": g_ev, "coef": coef, "train_loss": loss,
        "timing_s": tm,
        "error_small": {"shell_max": float(e_small_max.mean()), "shell_mean": float(e_small_mean.mean())},
        "error_cv": {"shell_max": float(e_cv_max.mean()), "shell_mean": float(e_cv_mean.mean())},
        "error_network": {"shell_max": float(e_g_max.mean()), "shell_mean": float(e_g_mean.mean())},
        "bias_cv": bias,
    }
