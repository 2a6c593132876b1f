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


def fq2_bound(rfq2, f0, tol=1e-5):
    """Allowed |fq2 - fq2_ref|.  fq2 is the mean over series of |a|^2 with
    a = <C(tau)>_tau; an error of at most tol |F(0)| in each a (the bound on
    fq) moves |a|^2 by at most 2 |a| tol |F(0)| + (tol |F(0)|)^2, i.e. fq2 by
    2 tol |F(0)| sqrt(fq2) + (tol F(0))^2; plus tol relative to fq2 itself."""
    f0 = abs(f0)
    return tol * abs(rfq2) + 2.0 * tol * f0 * abs(rfq2) ** 0.5 + (tol * f0) ** 2
