"""FP64 numpy restatement of the control-variate network (test
infrastructure): forward and backward of the MLP that libsassena_amd.so
trains on the GPU, with the same BF16 roundings of the MFMA operands
(inputs, W1, W2, hidden activations h1, and the back-propagated dz2 and
dz1), everything else in FP64.  ``grads_fd`` is the independent check of
the analytic backward: central finite differences of the unrounded loss."""
import numpy as np

IN, H = 16, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    xb = r(x)
    h1 = np.tanh(xb @ r(W1) + b1)
    h2 = np.tanh(r(h1) @ r(W2) + b2)
    y = h2 @ W3 + b3
    return y, (xb, h1, h2)


def grads(p, x, t, rounding=True):
    """Gradient of mean((y - t)^2) and the loss."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    y, (xb, h1, h2) = forward(p, x, rounding)
    n = y.shape[0]
    e = y - t
    dy = 2.0 * e / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, W3) * (1.0 - h2 * h2)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz2b = r(dz2)
    g[OFF_W2:OFF_B2] = (r(h1).T @ dz2b).ravel()
    dz1 = (dz2b @ r(W2).T) * (1.0 - h1 * h1)
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (xb.T @ r(dz1)).ravel()
    return g, float(np.mean(e * e))


def grads_fd(p, x, t, idx, h=1e-6):
    """Central differences of the unrounded loss for parameters idx."""
    out = []
    for i in idx:
        pp = np.array(p, dtype=np.float64)
        pm = pp.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((forward(pp, x, False)[0] - t) ** 2)
        lm = np.mean((forward(pm, x, False)[0] - t) ** 2)
        out.append((lp - lm) / (2 * h))
    return np.array(out)
