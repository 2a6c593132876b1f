"""Pins the FP64 oracle (oracle/sx_oracle.c) before it is trusted as the
checker of the HIP path: brute force, closed forms, invariants and the golden
vectors written by the independent numpy implementation in tests/golden/."""
import glob
import math
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def naive_dft(x, sign):
    n = x.size
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 13, 30, 97, 128, 400, 1000])
def test_fft_vs_naive_dft(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    for sign in (-1, 1):
        ref = naive_dft(x, sign)
        got = oracle_mod.fft(x, sign)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    # and against pocketfft at larger sizes
    assert np.allclose(oracle_mod.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-11 * n)


@pytest.mark.parametrize("NF", [1, 2, 9, 50, 201])
def test_fft_correlation_equals_direct_conjugate(oracle_mod, NF):
    """smath.cpp:141-156 and smath.cpp:51-76 give complex-conjugate results."""
    rng = np.random.default_rng(NF)
    a = rng.normal(size=NF) + 1j * rng.normal(size=NF)
    c_fft = oracle_mod.autocorrelate_fft(a)
    c_dir = oracle_mod.autocorrelate_direct(a)
    brute = np.array([np.sum(np.conj(a[:NF - t]) * a[t:]) / (NF - t) for t in range(NF)])
    assert np.allclose(c_fft, brute, rtol=0, atol=1e-12 * NF)
    assert np.allclose(c_dir, np.conj(brute), rtol=0, atol=1e-12 * NF)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_matches_golden(oracle_mod, path):
    g = np.load(path)
    kind = str(g["kind"])
    frames = g["frames"]
    coords = frames if kind == "all" else np.ascontiguousarray(frames.transpose(1, 0, 2))
    for tag, (dk, mt) in {"autocorrelate_fftw": ("autocorrelate", "fftw"),
                          "autocorrelate_direct": ("autocorrelate", "direct"),
                          "square_fftw": ("square", "fftw"),
                          "plain_fftw": ("plain", "fftw")}.items():
        for s, ql in enumerate(g["qlen"]):
            fqt, fq, fq2 = oracle_mod.compute(kind, coords, g["b"], ql * g["u"], dk, mt, nthreads=2)
            ref = g[f"fqt_{tag}"][s]
            scale = max(np.max(np.abs(ref)), 1e-300)
            assert np.max(np.abs(fqt - ref)) <= 1e-12 * scale, tag
            assert abs(fq - g[f"fq_{tag}"][s]) <= 1e-12 * scale, tag
            assert abs(fq2 - g[f"fq2_{tag}"][s]) <= 1e-12 * scale ** 2 + 1e-300, tag


def _traj(NA, NF, seed, box=20.0, step=0.2, static=False):
    rng = np.random.default_rng(seed)
    start = rng.uniform(0, box, size=(NA, 3))
    if static:
        frames = np.repeat(start[None], NF, axis=0)
    else:
        frames = start[None] + np.cumsum(rng.normal(0, step, size=(NF, NA, 3)), axis=0)
    return frames.astype(np.float32)


def test_static_atoms_give_constant_correlation(oracle_mod):
    frames = _traj(6, 20, 1, static=True)
    b = np.array([1.0, -2.0, 3.0, 0.5, 1.5, -0.7])
    q = np.array([[0.3, -1.1, 0.7], [1.2, 0.2, -0.4]])
    fqt, fq, fq2 = oracle_mod.compute("all", frames, b, q)
    A = [oracle_mod.scatter_all(frames, b, qq)[0] for qq in q]
    expect = np.mean([abs(x) ** 2 for x in A])
    assert np.allclose(fqt, expect, rtol=1e-12, atol=0)
    assert abs(fq - expect) <= 1e-12 * expect
    # self with one static atom: C(tau) = b^2 for every tau and vector
    one = frames[:, :1, :].transpose(1, 0, 2)
    fqt, fq, fq2 = oracle_mod.compute("self", one, b[:1], q)
    assert np.allclose(fqt, b[0] ** 2, rtol=1e-12, atol=0)
    assert abs(fq2 - b[0] ** 4) <= 1e-12 * b[0] ** 4


def test_single_atom_all_equals_self(oracle_mod):
    frames = _traj(1, 33, 2)
    b = np.array([2.5])
    q = np.array([[0.4, 0.5, -0.6], [-1.0, 0.1, 0.2], [0.0, 0.0, 2.0]])
    for dk in ("autocorrelate", "square", "plain"):
        a = oracle_mod.compute("all", frames, b, q, dk)
        s = oracle_mod.compute("self", frames.transpose(1, 0, 2), b, q, dk)
        assert np.allclose(a[0], s[0], rtol=0, atol=1e-12)
        assert abs(a[1] - s[1]) < 1e-12 and abs(a[2] - s[2]) < 1e-12
    # C(0) = |b|^2 for a single moving atom, whatever the path
    fqt, _, _ = oracle_mod.compute("all", frames, b, q, "autocorrelate")
    assert abs(fqt[0] - b[0] ** 2) < 1e-12


def test_self_is_linear_in_atoms(oracle_mod):
    frames = _traj(5, 16, 3)
    atoms = frames.transpose(1, 0, 2).copy()
    b = np.array([1.0, 2.0, -1.5, 0.3, 4.0])
    q = np.array([[0.7, 0.1, 0.2], [-0.3, 0.9, 0.4]])
    whole = oracle_mod.compute("self", atoms, b, q)
    parts = [oracle_mod.compute("self", atoms[j:j + 1], b[j:j + 1], q) for j in range(5)]
    assert np.allclose(whole[0], sum(p[0] for p in parts), rtol=1e-12, atol=1e-12)
    assert abs(whole[2] - sum(p[2] for p in parts)) < 1e-10


def test_square_and_plain_definitions(oracle_mod):
    frames = _traj(4, 10, 4)
    b = np.array([1.0, -0.5, 2.0, 0.25])
    q = np.array([[0.5, 0.5, 0.5], [1.0, -1.0, 0.0]])
    A = np.array([oracle_mod.scatter_all(frames, b, qq) for qq in q])
    fqt, fq, fq2 = oracle_mod.compute("all", frames, b, q, "square")
    assert np.allclose(fqt, np.mean(np.abs(A) ** 2, axis=0), rtol=1e-13, atol=0)
    fqt, fq, fq2 = oracle_mod.compute("all", frames, b, q, "plain")
    assert np.allclose(fqt, A.mean(axis=0), rtol=1e-13, atol=1e-14)
    a = A.mean(axis=1)
    assert abs(fq2 - np.mean(np.abs(a) ** 2)) < 1e-12


def test_gaussian_walk_self_closed_form(oracle_mod):
    """<C(tau)> = b^2 exp(-q^2 sigma^2 tau / 2) for a Gaussian walk."""
    NA, NF, sigma, ql = 400, 40, 0.1, 2.0
    rng = np.random.default_rng(5)
    frames = (rng.uniform(0, 10, size=(NA, 1, 3)) + np.cumsum(rng.normal(0, sigma, size=(NA, NF, 3)), axis=1)).astype(np.float32)
    u = rng.normal(size=(20, 3)); u /= np.linalg.norm(u, axis=1)[:, None]
    fqt, _, _ = oracle_mod.compute("self", frames, np.ones(NA), ql * u, nthreads=4)
    tau = np.arange(10)
    expect = NA * np.exp(-ql ** 2 * sigma ** 2 * tau / 2)
    assert np.allclose(fqt[:10].real, expect, rtol=0.03)
    assert np.max(np.abs(fqt[:10].imag)) < 0.03 * NA


def test_debye_limit_of_orientational_average(oracle_mod):
    """Static pair: the orientational mean of |A|^2 tends to the Debye sum."""
    r = np.array([[[0.0, 0.0, 0.0], [1.7, 2.2, -0.9]]], dtype=np.float32)
    b = np.array([1.3, -0.8])
    n = 4000  # Fibonacci sphere: a near-uniform deterministic set
    k = np.arange(n) + 0.5
    th = np.arccos(1 - 2 * k / n)
    ph = np.pi * (1 + 5 ** 0.5) * k
    u = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], axis=1)
    ql = 1.9
    d = np.linalg.norm(r[0, 1].astype(np.float64) - r[0, 0])
    debye = b[0] ** 2 + b[1] ** 2 + 2 * b[0] * b[1] * math.sin(ql * d) / (ql * d)
    fqt, fq, _ = oracle_mod.compute("all", np.repeat(r, 3, axis=0), b, ql * u, nthreads=4)
    assert abs(fq.real - debye) < 1e-4


def test_subvectors_and_scans(oracle_mod):
    q = np.array([0.3, 0.4, 1.2])
    u = oracle_mod.normalise_rows([[2.0, 0.0, 0.0], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    assert np.allclose(u[0], [1, 0, 0]) and np.all(u[1] == 0)
    assert np.allclose(np.linalg.norm(u[2]), 1.0)
    qv = oracle_mod.init_subvectors(q, u)
    assert np.allclose(qv, 1.3 * u)
    assert np.allclose(oracle_mod.init_subvectors(q, None), [q])
    assert len(oracle_mod.scan(0.1, 2.0, 0)) == 0
    assert np.allclose(oracle_mod.scan(0.1, 2.0, 1)[:, 0], [1.05])
    assert np.allclose(oracle_mod.scan(0.1, 2.0, 2)[:, 0], [0.1, 2.0])
    s = oracle_mod.scan(0.1, 2.0, 10, base=(0, 0, 1))[:, 2]
    j = np.arange(10)
    lin = 0.1 + (j / 9).astype(np.float32).astype(np.float64) * 1.9
    lin[-1] = 2.0
    assert np.array_equal(s, lin)
    s2 = oracle_mod.scan(0.0, 1.0, 5, exponent=2.0)[:, 0]
    frac = (np.arange(5) / 4).astype(np.float32)
    assert np.allclose(s2[1:4], (frac * frac)[1:4].astype(np.float64), rtol=0, atol=1e-7)


# ------------------------------------------------------------- multipole ---

def test_special_functions_match_scipy(oracle_mod):
    import scipy.special as sp
    for x in [1e-6, 0.3, 1.0, 4.2, 19.7, 20.3, 80.0, 866.0]:
        got = oracle_mod.sph_bessel_all(40, x)
        ref = sp.spherical_jn(np.arange(41), x)
        mask = np.abs(ref) > 1e-250
        assert np.allclose(got[mask], ref[mask], rtol=1e-11, atol=0)
    for n in [0, 1, 2, 7, 20, 39]:
        for x in [0.05, 1.0, 9.3, 45.0, 300.0]:
            assert abs(oracle_mod.cyl_bessel_j(n, x) - sp.jv(n, x)) < 1e-12 * max(1.0, abs(sp.jv(n, x))) + 1e-15
    for l in range(0, 16):
        for m in range(-l, l + 1):
            for th, ph in [(0.2, 0.7), (2.5, 4.0)]:
                assert abs(oracle_mod.ylm(l, m, th, ph) - sp.sph_harm_y(l, m, th, ph)) < 1e-13


def test_spherical_bessel_at_high_degree(oracle_mod):
    """j_l up to l = 200 (the largest accepted moment degree) against scipy,
    below and above the turning point x = l."""
    import scipy.special as sp
    for x in [0.5, 7.0, 30.0, 70.3, 150.0, 250.0]:
        got = oracle_mod.sph_bessel_all(200, x)
        ref = sp.spherical_jn(np.arange(201), x)
        mask = np.abs(ref) > 1e-250
        assert np.all(np.isfinite(got))
        assert np.allclose(got[mask], ref[mask], rtol=1e-11, atol=0)


def test_spherical_harmonics_at_high_degree(oracle_mod):
    """Y_lm up to the largest accepted degree (l = 200), where unnormalised
    associated Legendre values or (l+m)!/(l-m)! leave the double range:
    against scipy's sph_harm_y."""
    import scipy.special as sp
    for l, m in [(60, 5), (90, 88), (120, -100), (150, 151 - 1), (200, 3), (200, -199), (200, 200)]:
        for th, ph in [(0.3, 1.0), (1.2, 2.5), (2.9, 5.0)]:
            ref = sp.sph_harm_y(l, m, th, ph)
            got = oracle_mod.ylm(l, m, th, ph)
            assert abs(got - ref) <= 1e-10 * max(abs(ref), 1e-300) + 1e-300, (l, m, th, got, ref)


def test_sphere_multipole_reaches_debye_limit(oracle_mod):
    """(1/4 pi) sum_lm |A_lm|^2 of a static cluster tends to the Debye sum."""
    rng = np.random.default_rng(3)
    r = (rng.uniform(-1.5, 1.5, size=(1, 6, 3))).astype(np.float32)
    b = rng.uniform(-2, 3, 6)
    ql = 1.7
    mom = oracle_mod.sphere_moments(24)
    fqt, fq, fq2 = oracle_mod.compute_multipole("sphere", np.repeat(r, 2, axis=0), b, [0, 0, ql], mom)
    rr = r[0].astype(np.float64)
    d = np.linalg.norm(rr[:, None] - rr[None], axis=2)
    debye = np.sum(b[:, None] * b[None] * np.sinc(ql * d / np.pi))
    # the spherical coordinates are rounded to float, as the reference does
    assert abs(fqt[0].real - debye) < 2e-7 * abs(debye)
    assert abs(fq.real - debye) < 2e-7 * abs(debye)
    assert abs(fqt[0].imag) < 1e-12 * abs(debye)


def _moving_cluster(seed, NA=9, NF=3, lo=0.4, hi=3.5):
    """A few atoms near the origin, slightly moving, every coordinate in
    [lo, hi] (so z > 0 along a z-like axis)."""
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(lo, hi, size=(1, NA, 3))
    return (r0 + np.cumsum(rng.normal(0, 0.05, size=(NF, NA, 3)), axis=0)).astype(np.float32), rng.uniform(-2, 4, NA)


def test_sphere_moments_rebuild_the_plane_wave_amplitude(oracle_mod):
    """Plane-wave expansion e^{iq.r} = 4 pi sum_lm i^l j_l(qr) Y*_lm(r^) Y_lm(q^):
    the sphere moments A_lm(t) = 4 pi i^l sum_j b_j j_l(q r_j) Y*_lm(r^_j)
    (multipole_scatter_device.cpp:428-498) rebuild the plain amplitude
    A(q, t) = sum_j b_j e^{i q.r_j(t)} (all_vectors_scatter_device.cpp:418-438)
    at every direction q^ as sum_lm A_lm(t) Y_lm(q^).  A wrong power of i,
    a missing conjugate or a wrong sign of m breaks the identity (checked
    below on the moments themselves)."""
    frames, b = _moving_cluster(11, lo=-1.8, hi=1.8)
    ql, L = 1.6, 22
    A = {(l, m): oracle_mod.scatter_multipole("sphere", frames, b, [0, 0, ql], l, m)
         for l in range(L + 1) for m in range(-l, l + 1)}
    scale = np.sum(np.abs(b))
    for th, ph in [(0.3, 1.1), (1.9, 4.4), (2.8, 0.2)]:
        q = ql * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        direct = oracle_mod.scatter_all(frames, b, q)
        Y = {k: oracle_mod.ylm(k[0], k[1], th, ph) for k in A}
        rebuilt = sum(A[k] * Y[k] for k in A)
        # the reference rounds r, theta, phi to float: ~1e-7 of sum |b|
        assert np.max(np.abs(rebuilt - direct)) < 2e-6 * scale
        # planted mistakes: conj(Y) on the q side, or i^-l instead of i^l
        wrong1 = sum(A[k] * np.conj(Y[k]) for k in A)
        wrong2 = sum(A[k] * Y[k] * (-1) ** k[0] for k in A)
        assert np.max(np.abs(wrong1 - direct)) > 1e-3 * scale
        assert np.max(np.abs(wrong2 - direct)) > 1e-3 * scale


@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.3, -0.2, 0.9)])
def test_cylinder_moments_rebuild_the_plane_wave_amplitude(oracle_mod, axis):
    """Jacobi-Anger in the cylinder frame of the axis (e_r, e_phi, e_z of
    coor3d.cpp:278-296): e^{i q.r} = e^{i q_z z} [J_0(q_r rho) + 2 sum_n i^n
    J_n(q_r rho) cos(n (phi - phi'))] for q at azimuth phi'.  With the
    reference's moments of q (azimuth phi_q, multipole_scatter_device.cpp:
    900-987: m = 0, 1 the cos / sin parts of order 2l with (-1)^l, m = 2, 3
    those of order 2l - 1 with i (-1)^(l-1), all times sqrt(2 pi) sqrt(2)),
    the plain amplitude at the rotated vector q' (same q_r, q_z, azimuth
    phi_q + D) is
      A(q', t) = [A_00 + sqrt(2) sum_l (A_l0 cos 2lD + A_l1 sin 2lD
                  + A_l2 cos (2l-1)D + A_l3 sin (2l-1)D)] / sqrt(2 pi).
    Every atom has z > 0 and q_z > 0 (parallel_sign +1).  The check pins each
    of the four moment kinds, their signs, orders and the frame; mistakes
    planted in the combination (m = 2 and 3 swapped, (-1)^l dropped, the
    sign of phi - phi_q flipped) fail it."""
    frames, b = _moving_cluster(12 if axis[2] == 1.0 else 13)
    base = oracle_mod.cylinder_base(axis)
    er, ep, ez = base
    qr, qz, phq = 1.3, 0.8, 0.9
    q = qr * (np.cos(phq) * er + np.sin(phq) * ep) + qz * ez
    L = 18
    mom = [(0, 0)] + [(l, m) for l in range(1, L + 1) for m in range(4)]
    A = {k: oracle_mod.scatter_multipole("cylinder", frames, b, q, k[0], k[1], axis=axis) for k in mom}
    scale = np.sum(np.abs(b))

    def rebuild(D, swap=False, nosign=False, flip=False):
        tot = A[(0, 0)].copy()
        for l in range(1, L + 1):
            s = (-1) ** l if nosign else 1
            d = -D if flip else D
            a2, a3 = (A[(l, 3)], A[(l, 2)]) if swap else (A[(l, 2)], A[(l, 3)])
            tot += np.sqrt(2) * (s * A[(l, 0)] * np.cos(2 * l * d) + s * A[(l, 1)] * np.sin(2 * l * d)
                                 + a2 * np.cos((2 * l - 1) * d) + a3 * np.sin((2 * l - 1) * d))
        return tot / np.sqrt(2 * np.pi)

    for D in (0.0, 0.7, 2.3, 4.1):
        qd = qr * (np.cos(phq + D) * er + np.sin(phq + D) * ep) + qz * ez
        direct = oracle_mod.scatter_all(frames, b, qd)
        assert np.max(np.abs(rebuild(D) - direct)) < 2e-6 * scale, D
        if D != 0.0:
            for bad in ({"swap": True}, {"nosign": True}, {"flip": True}):
                assert np.max(np.abs(rebuild(D, **bad) - direct)) > 1e-3 * scale, (D, bad)


@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.3, -0.2, 0.9)])
def test_cylinder_multipole_sum_is_grafs_theorem(oracle_mod, axis):
    """Static cluster: (1 / 2 pi) sum over l <= L, m = 0..3 of |A_lm|^2, the
    cylinder device's fqt[0] (multipole_scatter_device.cpp:866), is the
    average of |A(q)|^2 over the azimuth of q about the axis, which Graf's
    addition theorem gives in closed form:
      sum_ij b_i b_j e^{i q_z (z_i - z_j)} J_0(q_r |rho_i - rho_j|)."""
    import scipy.special as sp
    frames, b = _moving_cluster(14, NA=7, NF=1)
    frames = np.repeat(frames, 2, axis=0)
    base = oracle_mod.cylinder_base(axis)
    er, ep, ez = base
    qr, qz = 1.1, 0.6
    q = qr * (np.cos(0.4) * er + np.sin(0.4) * ep) + qz * ez
    fqt, fq, fq2 = oracle_mod.compute_multipole("cylinder", frames, b, q, oracle_mod.cylinder_moments(30), axis=axis)
    r = frames[0].astype(np.float64)
    z = r @ ez
    rho = r - z[:, None] * ez[None]
    dr = np.linalg.norm(rho[:, None] - rho[None], axis=2)
    graf = np.sum(b[:, None] * b[None] * np.exp(1j * qz * (z[:, None] - z[None])) * sp.jv(0, qr * dr))
    assert abs(fqt[0] - graf) < 1e-6 * np.sum(np.abs(b)) ** 2
    assert abs(fqt[1] - graf) < 1e-6 * np.sum(np.abs(b)) ** 2
