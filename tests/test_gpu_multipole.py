"""Parity of the multipole orientational averages (sx_compute_multipole,
through the C ABI) with the FP64 oracle.  Same bound as the vector
averages: max_tau |F - F_ref| / |F_ref(0)| < 1e-5 per shell."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import fq2_bound, shell_error  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-5


def centred_walk(NA, NF, seed, half=8.0, step=0.15):
    """Seeded random walk about the origin (the multipole expansions are
    about the coordinate origin), frames-major [NF][NA][3] float32."""
    rng = np.random.default_rng(seed)
    start = rng.uniform(-half, half, size=(NA, 3))
    frames = start[None] + np.cumsum(rng.normal(0, step, size=(NF, NA, 3)), axis=0)
    b = rng.choice([-3.74, 6.65, 9.36, 5.80], size=NA)
    return frames.astype(np.float32), b


def check(fqt, fq, fq2, ref):
    rfqt, rfq, rfq2 = ref
    scale = abs(rfqt[0])
    err = shell_error(fqt, rfqt)[0]
    assert err < TOL, f"fqt shell error {err:.3e}"
    assert abs(fq - rfq) <= TOL * scale, (fq, rfq)
    assert abs(fq2 - rfq2) <= fq2_bound(rfq2, scale, TOL), (fq2, rfq2)


def run(kind, frames, b, qv, moments, axis=(0.0, 0.0, 1.0), dsp="autocorrelate", method="fftw", b_shell=None,
        atom_split=0, terms=0):
    import sassena_amd
    with sassena_amd.ScatterDevice("all", dsp, method, atom_split=atom_split, multipole_terms=terms) as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        return dev.compute_multipole(kind, qv, moments, axis=axis, b_shell=b_shell)


@pytest.mark.parametrize("terms", [1, 2, 3])
@pytest.mark.parametrize("dsp,method", [("autocorrelate", "fftw"), ("autocorrelate", "direct"), ("square", "fftw"),
                                        ("plain", "fftw")])
def test_sphere_against_oracle(oracle_mod, dsp, method, terms):
    """Every source of the spherical coordinates and Bessel functions: the
    per-(atom, frame) pre-pass with the table of j_l per (shell, atom,
    frame) (1), the kernel forming coordinates and Bessel functions itself
    (2, the path of trajectories whose terms exceed 16 GB), and the pre-pass
    with j_0, j_1 only, the higher degrees by recurrence in the kernel (3,
    the path of tables beyond their budget)."""
    frames, b = centred_walk(120, 257, 11)
    qv = np.array([[0.0, 0.0, 0.3], [0.5, 0.2, -0.7], [1.6, 0.4, 0.1]])
    mom = oracle_mod.sphere_moments(7)
    fqt, fq, fq2 = run("sphere", frames, b, qv, mom, dsp=dsp, method=method, terms=terms)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("sphere", frames, b, qv[s], mom, dsp=dsp,
                                                                   method=method))


@pytest.mark.parametrize("terms", [1, 2, 3])
@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.3, -0.2, 0.9), (1.0, 0.0, 0.0)])
def test_cylinder_against_oracle(oracle_mod, axis, terms):
    """Every source of the cylinder-frame coordinates and Bessel functions:
    the per-(atom, frame) pre-pass with the table of J_n per (shell, atom,
    frame) (1), the kernel forming them itself (2), and the pre-pass with
    J_0, J_1 only, the higher orders by recurrence in the kernel (3)."""
    frames, b = centred_walk(100, 300, 12)
    qv = np.array([[0.4, 0.1, 0.2], [-0.3, 0.9, -0.5], [0.0, 0.0, 1.1], [1.2, -0.6, 0.0]])
    mom = oracle_mod.cylinder_moments(6)
    fqt, fq, fq2 = run("cylinder", frames, b, qv, mom, axis=axis, terms=terms)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("cylinder", frames, b, qv[s], mom, axis=axis))


@pytest.mark.parametrize("terms", [1, 2])
def test_sphere_high_degrees(oracle_mod, terms):
    """Degrees up to the accepted maximum of 200, where unnormalised
    associated Legendre values and (l+m)!/(l-m)! leave the double range: the
    recurrence starts from the m-normalised value and the remaining factor
    comes from log-gamma, so every moment stays finite; against the oracle
    (pinned to scipy at these degrees)."""
    frames, b = centred_walk(40, 70, 16, half=6.0)
    qv = np.array([[0.0, 0.0, 7.0], [3.0, -4.0, 2.0]])   # |q| r up to ~70: j_l(q r) alive to l ~ 80
    mom = [(l, m) for l in (60, 90, 120, 150, 200) for m in (0, l // 2, -(l - 1), l)]
    fqt, fq, fq2 = run("sphere", frames, b, qv, mom, terms=terms)
    assert np.all(np.isfinite(fqt)) and np.all(np.isfinite(fq)) and np.all(np.isfinite(fq2))
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("sphere", frames, b, qv[s], mom))


@pytest.mark.parametrize("lmax", [16, 17])
def test_sphere_bessel_table_degrees(oracle_mod, lmax):
    """The Bessel table holds degrees up to 16: at 16 every moment reads it,
    at 17 the call falls back to j_0, j_1 and the recurrence in the kernel.
    |q| r spans both branches of the recurrence (upward above the degree,
    Miller's below), and a shell of |q| = 0 gives j_l(0)."""
    frames, b = centred_walk(60, 90, 21)
    qv = np.array([[0.0, 0.0, 0.0], [0.2, -0.1, 0.3], [1.5, 0.9, -0.4], [3.0, 1.0, 2.0]])
    mom = [(l, m) for l in (0, 1, 2, 5, 9, 12, lmax) for m in sorted({0, 1, -1, l // 2, -l, l}) if -l <= m <= l]
    fqt, fq, fq2 = run("sphere", frames, b, qv, mom, terms=1)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("sphere", frames, b, qv[s], mom))


@pytest.mark.parametrize("lmax", [16, 17])
def test_cylinder_bessel_table_orders(oracle_mod, lmax):
    """The cylinder's Bessel table holds orders up to 32 (moments up to
    l = 16): at 16 every moment reads it, at 17 (order 34) the call falls
    back to J_0, J_1 and the recurrence in the kernel.  q_r r spans the
    upward and Miller branches, and a shell along the axis has q_r = 0."""
    frames, b = centred_walk(60, 90, 22)
    qv = np.array([[0.0, 0.0, 0.8], [0.3, -0.2, 0.1], [1.4, 0.7, -0.3], [3.1, 0.8, 1.5]])
    mom = [(0, 0)] + [(l, m) for l in (1, 2, 6, 11, lmax) for m in range(4)]
    fqt, fq, fq2 = run("cylinder", frames, b, qv, mom, terms=1)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("cylinder", frames, b, qv[s], mom))


def test_large_moment_lists_and_per_shell_factors(oracle_mod):
    """More moments than one moment block (8) and one block group (32),
    high orders, ragged frame count and per-shell scattering factors."""
    frames, b = centred_walk(70, 131, 13)
    qv = np.array([[0.9, 0.0, 0.0], [0.0, 2.2, 0.5]])
    mom = oracle_mod.sphere_moments(14)  # 225 moments, l up to 14
    rng = np.random.default_rng(5)
    bs = rng.uniform(-2.0, 4.0, size=(len(qv), frames.shape[1]))
    fqt, fq, fq2 = run("sphere", frames, b, qv, mom, b_shell=bs)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("sphere", frames, bs[s], qv[s], mom))


@pytest.mark.parametrize("kind", ["sphere", "cylinder"])
@pytest.mark.parametrize("split", [1, 5])
def test_atom_ranges(oracle_mod, kind, split):
    """The atoms summed in ranges whose partial amplitudes the correlation
    adds: one range, and five ragged ones (256 padded atoms, 52 a range),
    with per-shell factors and the square and autocorrelation transforms."""
    frames, b = centred_walk(230, 150, 17)
    qv = np.array([[0.2, 0.6, -0.1], [1.3, -0.4, 0.8]])
    mom = oracle_mod.sphere_moments(5) if kind == "sphere" else oracle_mod.cylinder_moments(4)
    rng = np.random.default_rng(8)
    bs = rng.uniform(-2.0, 4.0, size=(len(qv), frames.shape[1]))
    for dsp in ("autocorrelate", "square"):
        for terms in ((1, 2) if kind == "sphere" else (0,)):
            fqt, fq, fq2 = run(kind, frames, b, qv, mom, dsp=dsp, b_shell=bs, atom_split=split, terms=terms)
            for s in range(len(qv)):
                check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole(kind, frames, bs[s], qv[s], mom,
                                                                           dsp=dsp))


def test_unsorted_moment_list(oracle_mod):
    """Moments in no particular order (the Bessel recurrence then compares
    every step with every degree instead of running in segments), small and
    large |q| r (Miller's and the upward recurrence)."""
    frames, b = centred_walk(90, 140, 14)
    qv = np.array([[0.0, 0.05, 0.0], [0.7, -0.4, 0.3], [2.5, 0.0, 1.0]])
    mom = oracle_mod.sphere_moments(9)
    order = np.random.default_rng(3).permutation(len(mom))
    mom = [mom[i] for i in order]
    fqt, fq, fq2 = run("sphere", frames, b, qv, mom)
    for s in range(len(qv)):
        check(fqt[s], fq[s], fq2[s], oracle_mod.compute_multipole("sphere", frames, b, qv[s], mom))


def test_sphere_converges_to_vector_average(oracle_mod):
    """A static cluster: the truncated multipole sum and a dense vector
    average both tend to the Debye sum."""
    rng = np.random.default_rng(8)
    r = rng.uniform(-1.2, 1.2, size=(1, 5, 3)).astype(np.float32)
    frames = np.repeat(r, 4, axis=0)
    b = rng.uniform(0.5, 3.0, 5)
    ql = 2.1
    fqt, fq, _ = run("sphere", frames, b, np.array([[0.0, ql, 0.0]]), oracle_mod.sphere_moments(22))
    rr = r[0].astype(np.float64)
    d = np.linalg.norm(rr[:, None] - rr[None], axis=2)
    debye = np.sum(b[:, None] * b[None] * np.sinc(ql * d / np.pi))
    assert abs(fq[0].real - debye) < 2e-6 * debye


def test_multipole_errors():
    import sassena_amd
    frames, b = centred_walk(10, 20, 1)
    with sassena_amd.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        with pytest.raises(sassena_amd.SxError):
            dev.compute_multipole("sphere", [[0.1, 0.0, 0.0]], [[2, 3]])
        with pytest.raises(sassena_amd.SxError):
            dev.compute_multipole("cylinder", [[0.1, 0.0, 0.0]], [[1, 4]])
        with pytest.raises(sassena_amd.SxError):
            dev.compute_multipole("cylinder", [[0.1, 0.0, 0.0]], [[0, 1]])
        with pytest.raises(sassena_amd.SxError):
            dev.compute_multipole("cylinder", [[0.1, 0.0, 0.0]], [[0, 0]], axis=(0.0, 0.0, 0.0))
    with sassena_amd.ScatterDevice("self") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        with pytest.raises(sassena_amd.SxError):
            dev.compute_multipole("sphere", [[0.1, 0.0, 0.0]], [[0, 0]])
