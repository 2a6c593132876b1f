"""Learned control variate for the orientational average.

The plain estimate of a coherent shell average uses NM orientations (1,000 in
the large case); its cost is linear in NM.  Here a small network, trained on
the GPU by libsassena_amd.so's own MFMA kernels, maps cheap features of a
shell and lag -- |q|, log lag, and the 16-orientation estimate F16(q, tau) --
to the many-orientation average.  It is trained on the first part of the
trajectory, where both estimates are computed, and then corrects the
16-orientation estimate on the rest of the trajectory.  The plain
many-orientation average stays the fallback and is what the result is judged
against (``evaluate``).

All values are normalised by nb = sum_j b_j^2 (the large-|q| limit of
F(q, 0) for uncorrelated positions), which is known exactly from the
scattering factors.
"""
from __future__ import annotations

import time

import numpy as np

from .device import ScatterDevice
from .mlp import MLP

NFEAT = 16


def features(qlen, f16, nb, tmax, qmax):
    """Rows (shell, lag) of the network input: [|q|/qmax, log1p(tau)/log1p(tmax),
    tau/tmax, Re F16/nb, Im F16/nb, Re F16(q,0)/nb, (|q|/qmax)^2 log1p(tau)/log1p(tmax)],
    zero padded to 16 columns."""
    ns, NF = f16.shape
    tau = np.arange(NF, dtype=np.float64)
    q = np.repeat(np.asarray(qlen, np.float64) / qmax, NF)
    lt = np.tile(np.log1p(tau) / np.log1p(tmax), ns)
    x = np.zeros((ns * NF, NFEAT), np.float32)
    x[:, 0] = q
    x[:, 1] = lt
    x[:, 2] = np.tile(tau / tmax, ns)
    x[:, 3] = (f16.real / nb).ravel()
    x[:, 4] = (f16.imag / nb).ravel()
    x[:, 5] = np.repeat(f16[:, 0].real / nb, NF)
    x[:, 6] = q * q * lt
    return x


def shell_errors(f, ref):
    """Per shell: max and mean over lags of |f - ref| / |ref(0)|."""
    d = np.abs(f - ref) / np.abs(ref[:, :1])
    return d.max(axis=1), d.mean(axis=1)


def run(xyz, b, qlen, u, ntrain_frames, n_small=16, steps=3000, batch=4096, lr=2e-3, seed=1, device=0):
    """The full pipeline on one trajectory (xyz [3][NA][NF] float32).

    Returns a dict with the estimates on the evaluation frames (the frames
    after ntrain_frames), the many-orientation reference there, timings
    (seconds) of each stage and the errors of both estimates."""
    qlen = np.asarray(qlen, np.float64)
    u = np.asarray(u, np.float64)
    nb = float(np.sum(np.asarray(b, np.float64) ** 2))
    NF = xyz.shape[2]
    tmax = float(NF)
    qmax = float(np.max(qlen))
    train = np.ascontiguousarray(xyz[:, :, :ntrain_frames])
    evalf = np.ascontiguousarray(xyz[:, :, ntrain_frames:])
    usmall = u[:n_small]
    tm = {}

    def shells(seg, orient):
        with ScatterDevice("all", device=device) as dev:
            dev.stage(seg, layout="soa")
            dev.set_factors(b)
            t0 = time.perf_counter()
            fqt, _, _ = dev.compute_oriented(qlen, orient)
            return fqt, time.perf_counter() - t0

    # training data: both estimates on the first frames
    f_many_tr, tm["many_train"] = shells(train, u)
    f_small_tr, tm["small_train"] = shells(train, usmall)
    x_tr = features(qlen, f_small_tr, nb, tmax, qmax)
    y_tr = (f_many_tr.real / nb).ravel().astype(np.float32)
    net = MLP(device=device, seed=seed)
    t0 = time.perf_counter()
    loss = net.train(x_tr, y_tr, steps=steps, batch=batch, lr=lr, seed=seed)
    tm["train"] = time.perf_counter() - t0

    # evaluation frames: the cheap estimate and its correction
    f_small_ev, tm["small_eval"] = shells(evalf, usmall)
    x_ev = features(qlen, f_small_ev, nb, tmax, qmax)
    t0 = time.perf_counter()
    y_ev = net.predict(x_ev).reshape(f_small_ev.shape).astype(np.float64)
    tm["predict"] = time.perf_counter() - t0
    f_cv = nb * y_ev + 1j * f_small_ev.imag
    # the judge: the plain many-orientation average on the same frames
    f_many_ev, tm["many_eval"] = shells(evalf, u)
    net.close()

    e_small_max, e_small_mean = shell_errors(f_small_ev, f_many_ev)
    e_cv_max, e_cv_mean = shell_errors(f_cv, f_many_ev)
    bias = float(np.mean((f_cv.real - f_many_ev.real) / np.abs(f_many_ev[:, :1].real)))
    return {
        "f_cv": f_cv, "f_small": f_small_ev, "f_many": f_many_ev, "train_loss": loss, "timing_s": tm,
        "error_small": {"shell_max": float(e_small_max.mean()), "shell_mean": float(e_small_mean.mean())},
        "error_cv": {"shell_max": float(e_cv_max.mean()), "shell_mean": float(e_cv_mean.mean())},
        "bias_cv": bias,
    }
