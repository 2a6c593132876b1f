"""Shared helpers of the parity tests."""
import numpy as np


def shell_error(fqt, ref):
    """max_tau |F(tau) - F_ref(tau)| / |F_ref(0)| per shell (rows)."""
    fqt = np.atleast_2d(fqt)
    ref = np.atleast_2d(ref)
    return np.max(np.abs(fqt - ref), axis=1) / np.abs(ref[:, 0])


def small_trajectory(NA, NF, seed, box=20.0, step=0.15):
    """Seeded random walk, frames-major [NF][NA][3] float32."""
    rng = np.random.default_rng(seed)
    start = rng.uniform(0, box, size=(NA, 3))
    frames = start[None] + np.cumsum(rng.normal(0, step, size=(NF, NA, 3)), axis=0)
    b = rng.choice([-3.74, 6.65, 9.36, 5.80], size=NA)
    return frames.astype(np.float32), b


def unit_vectors(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1)[:, None]
