"""Learned control variate for the orientational average, with the plain
many-orientation average as fallback.

The plain estimate of a coherent shell average uses NM orientations (1,000 in
the large case); its cost is linear in NM.  A cheap estimate F16 uses the
first 16 of them.  Where the orientational sampling error of F16 persists in
time -- the anisotropy of a crystal, of the simulation box at small |q| --
the correction D = F_many - F16 measured on early frames predicts the
correction on later ones, and a small network learns how much of it carries
over, per shell and lag.

Training (the first ``ntrain_frames`` frames, split in halves A and B; both
estimates are computed on each half):

    x = features(F16_B, D_A),  y = D_B / s_q

where s_q is the shell scale |F_many(q, 0)| of the training halves.  The
network is trained by libsassena_amd.so's own MFMA kernels.  It is then
validated out of sample in the reverse direction (features of F16_A and D_B,
against D_A): a shell whose validation error exceeds ``tol`` falls back to
the plain many-orientation average on the evaluation frames.

Evaluation (the remaining frames): F_cv = F16 + s_q g(F16_E, D_B) on the
shells that passed, the plain average elsewhere.  ``run`` also computes the
plain many-orientation average on every shell of the evaluation frames, the
referee of the error, and times each stage.
"""
from __future__ import annotations

import time

import numpy as np

from .device import ScatterDevice
from .mlp import MLP

NFEAT = 16


def features(qlen, f16, dprev, scale, nb, tmax, qmax):
    """Rows (shell, lag), zero padded to 16 columns: |q|/qmax, log1p(tc)/
    log1p(tmax), tc/tmax (tc = lag clamped to tmax - 1), Re and Im F16/s,
    Re F16(q,0)/s, Re and Im D_prev(q, tc)/s, Re D_prev(q, 0)/s and
    log10(s/nb)/4, with s the shell scale."""
    ns, NF = f16.shape
    tcl = np.minimum(np.arange(NF), int(tmax) - 1)
    s = np.asarray(scale, np.float64)[:, None]
    dp = dprev[:, tcl] / s
    x = np.zeros((ns * NF, NFEAT), np.float32)
    x[:, 0] = np.repeat(np.asarray(qlen, np.float64) / qmax, NF)
    x[:, 1] = np.tile(np.log1p(tcl) / np.log1p(tmax), ns)
    x[:, 2] = np.tile(tcl / tmax, ns)
    x[:, 3] = (f16.real / s).ravel()
    x[:, 4] = (f16.imag / s).ravel()
    x[:, 5] = np.repeat(f16[:, 0].real / s[:, 0], NF)
    x[:, 6] = dp.real.ravel()
    x[:, 7] = dp.imag.ravel()
    x[:, 8] = np.repeat(dprev[:, 0].real / s[:, 0], NF)
    x[:, 9] = np.repeat(np.log10(np.maximum(s[:, 0], 1e-300) / nb) / 4.0, NF)
    return x


def shell_errors(f, ref):
    """Per shell: max and mean over lags of |f - ref| / |ref(0)|."""
    d = np.abs(f - ref) / np.abs(ref[:, :1])
    return d.max(axis=1), d.mean(axis=1)


def referee_sampling_error(xyz, b, qlen, u, device=0):
    """The orientational sampling error of the many-orientation average
    itself, by a split half: the two halves of the orientation set give
    averages F_a, F_b whose difference has variance 4 / NM of the
    per-orientation spread, so |F_a - F_b| / 2 estimates the error of the full
    NM-orientation average.  Per shell: the mean over lags of that relative
    to |F(q, 0)|."""
    u = np.asarray(u, np.float64)
    h = u.shape[0] // 2
    with ScatterDevice("all", device=device) as dev:
        dev.stage(np.ascontiguousarray(xyz), layout="soa")
        dev.set_factors(b)
        fa, _, _ = dev.compute_oriented(qlen, u[:h])
        fb, _, _ = dev.compute_oriented(qlen, u[h:2 * h])
    f = 0.5 * (fa + fb)
    return np.mean(np.abs(fa - fb) / 2.0 / np.abs(f[:, :1]), axis=1)


def run(xyz, b, qlen, u, ntrain_frames, n_small=16, tol=1e-3, steps=3000, batch=4096, lr=2e-3, seed=1,
        device=0, progress=None):
    """The full pipeline on one trajectory (xyz [3][NA][NF] float32).

    Returns the estimates on the evaluation frames (the frames after
    ntrain_frames) -- f_cv (learned control variate with fallback), f_net
    (control variate on every shell), f_small, f_many (the referee) -- the
    shells that use the control variate, timings (seconds) of each stage and
    the errors.  ``progress(stage, seconds)`` is called after each stage."""
    note = progress or (lambda stage, sec: None)
    qlen = np.asarray(qlen, np.float64)
    u = np.asarray(u, np.float64)
    nb = float(np.sum(np.asarray(b, np.float64) ** 2))
    qmax = float(np.max(qlen))
    half = ntrain_frames // 2
    seg_a = np.ascontiguousarray(xyz[:, :, :half])
    seg_b = np.ascontiguousarray(xyz[:, :, half:2 * half])
    evalf = np.ascontiguousarray(xyz[:, :, ntrain_frames:])
    tmax = float(half)
    usmall = u[:n_small]
    tm = {"many_train": 0.0, "small_train": 0.0}

    def shells(seg, orient, sel=None):
        q = qlen if sel is None else qlen[sel]
        with ScatterDevice("all", device=device) as dev:
            dev.stage(seg, layout="soa")
            dev.set_factors(b)
            t0 = time.perf_counter()
            fqt, _, _ = dev.compute_oriented(q, orient)
            return fqt, time.perf_counter() - t0

    f_many = {}
    f_small = {}
    for name, seg in (("A", seg_a), ("B", seg_b)):
        f_many[name], t = shells(seg, u)
        tm["many_train"] += t
        note(f"many-orientation average, training half {name}", t)
        f_small[name], t = shells(seg, usmall)
        tm["small_train"] += t
    d_a = f_many["A"] - f_small["A"]
    d_b = f_many["B"] - f_small["B"]
    scale = 0.5 * (np.abs(f_many["A"][:, 0]) + np.abs(f_many["B"][:, 0]))
    scale = np.where(scale > 0, scale, nb)

    net = MLP(device=device, seed=seed)
    x_tr = features(qlen, f_small["B"], d_a, scale, nb, tmax, qmax)
    y_tr = (d_b.real / scale[:, None]).ravel().astype(np.float32)
    t0 = time.perf_counter()
    loss = net.train(x_tr, y_tr, steps=steps, batch=batch, lr=lr, seed=seed)
    tm["train"] = time.perf_counter() - t0
    note("network training", tm["train"])
    # out-of-sample validation in the reverse direction, per shell
    g_val = net.predict(features(qlen, f_small["A"], d_b, scale, nb, tmax, qmax)).reshape(d_a.shape)
    f_val = f_small["A"] + scale[:, None] * g_val
    _, e_val = shell_errors(f_val, f_many["A"])
    use_cv = e_val <= tol

    # evaluation frames: the cheap estimate, its correction, the fallback shells
    f_small_ev, tm["small_eval"] = shells(evalf, usmall)
    note("16-orientation estimate, evaluation frames", tm["small_eval"])
    t0 = time.perf_counter()
    g_ev = net.predict(features(qlen, f_small_ev, d_b, scale, nb, tmax, qmax)).reshape(f_small_ev.shape)
    f_net = f_small_ev + scale[:, None] * g_ev.astype(np.float64)
    tm["predict"] = time.perf_counter() - t0
    f_cv = f_net.copy()
    tm["fallback"] = 0.0
    if not np.all(use_cv):
        fb, tm["fallback"] = shells(evalf, u, np.flatnonzero(~use_cv))
        f_cv[~use_cv] = fb
        note(f"fallback on {int(np.count_nonzero(~use_cv))} shells", tm["fallback"])
    # the referee: the plain many-orientation average on every shell
    f_many_ev, tm["many_eval"] = shells(evalf, u)
    note("many-orientation average (referee), evaluation frames", tm["many_eval"])
    net.close()

    def summary(f):
        emax, emean = shell_errors(f, f_many_ev)
        return {"shell_max": float(emax.mean()), "shell_mean": float(emean.mean())}

    bias = float(np.mean((f_net.real - f_many_ev.real) / np.abs(f_many_ev[:, :1].real)))
    cv_time = tm["small_eval"] + tm["predict"] + tm["fallback"]
    # How much of the 16-orientation error persists from one training half to
    # the other: per shell, the correlation over lags of Re D_A and Re D_B (a
    # crystal's anisotropy: near 1; a diffusing gas: near 0 beyond the box
    # form factor), and the part of D_B that D_A explains.
    persistence = np.zeros(len(qlen))
    explained = np.zeros(len(qlen))
    for s_ in range(len(qlen)):
        da, db = d_a[s_].real, d_b[s_].real
        if np.std(da) > 0 and np.std(db) > 0:
            persistence[s_] = float(np.corrcoef(da, db)[0, 1])
        explained[s_] = 1.0 - float(np.sum((db - da) ** 2) / max(np.sum(db ** 2), 1e-300))
    return {
        "f_cv": f_cv, "f_net": f_net, "f_small": f_small_ev, "f_many": f_many_ev, "use_cv": use_cv,
        "validation_error": e_val, "train_loss": loss, "timing_s": tm,
        "error_small": summary(f_small_ev), "error_net": summary(f_net), "error_cv": summary(f_cv),
        "bias_net": bias, "cv_shells": int(np.count_nonzero(use_cv)),
        "persistence": persistence, "explained": explained,
        "speedup_eval": tm["many_eval"] / cv_time if cv_time > 0 else None,
    }
