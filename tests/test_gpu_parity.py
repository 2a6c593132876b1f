"""Parity of the gfx950 path (through the C ABI) with the FP64 oracle and the
golden fixtures.  Tolerance: the per-shell error max_tau |F - F_ref| /
|F_ref(0)| must stay below 1e-5 (FP32 sincos, FP32 partial sums folded into FP64 and FP32 FFT against
FP64 references); fq and fq2 are held to the same relative bound."""
import glob
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import fq2_bound, shell_error, small_trajectory, unit_vectors  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.fixture(scope="module")
def sx():
    import sassena_amd
    return sassena_amd


def check_close(fqt, fq, fq2, rfqt, rfq, rfq2, tol=TOL):
    """One shell: fqt by the per-shell error, fq against |F_ref(0)|, fq2 by
    the bound the fq bound implies (util.fq2_bound)."""
    scale = abs(rfqt[0])
    err = shell_error(fqt, rfqt)[0]
    assert err < tol, f"fqt shell error {err:.3e}"
    assert abs(fq - rfq) <= tol * scale, (fq, rfq)
    assert abs(fq2 - rfq2) <= fq2_bound(rfq2, scale, tol), (fq2, rfq2)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
@pytest.mark.parametrize("tag", ["autocorrelate_fftw", "autocorrelate_direct", "square_fftw", "plain_fftw"])
def test_golden(sx, path, tag):
    g = np.load(path)
    kind = str(g["kind"])
    dsp, method = tag.rsplit("_", 1)
    qv = g["qlen"][:, None, None] * g["u"][None]
    with sx.ScatterDevice(kind, dsp, method) as dev:
        dev.stage(g["frames"], layout="frames")
        dev.set_factors(g["b"])
        fqt, fq, fq2 = dev.compute_shells(qv)
    ref = g[f"fqt_{tag}"]
    for s in range(len(g["qlen"])):
        check_close(fqt[s], fq[s], fq2[s], ref[s], g[f"fq_{tag}"][s], g[f"fq2_{tag}"][s])


@pytest.mark.parametrize("oriented", [False, True])
@pytest.mark.parametrize("dsp", ["autocorrelate", "square", "plain"])
def test_c1_all_against_oracle(sx, oracle_mod, dsp, oriented):
    """The full CPU-runnable case: 1000 atoms, 200 frames, 10 shells x 50 vectors."""
    from sassena_amd import synth
    T = synth.case("c1")
    frames = synth.to_frames_aos(T.xyz)
    qv = np.array([T.qvectors(s) for s in range(T.nshells)])
    with sx.ScatterDevice("all", dsp) as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        if oriented:
            fqt, fq, fq2 = dev.compute_oriented(T.qlen, T.orient)
        else:
            fqt, fq, fq2 = dev.compute_shells(qv)
    for s in range(T.nshells):
        r = oracle_mod.compute("all", frames, T.b, qv[s], dsp, nthreads=8)
        check_close(fqt[s], fq[s], fq2[s], *r)


@pytest.mark.parametrize("dsp", ["autocorrelate", "square", "plain"])
def test_self_against_oracle(sx, oracle_mod, dsp):
    frames, b = small_trajectory(120, 300, seed=7, box=30.0, step=0.2)
    atoms = np.ascontiguousarray(frames.transpose(1, 0, 2))
    u = unit_vectors(12, 8)
    qlen = np.array([0.3, 1.1, 2.9])
    qv = qlen[:, None, None] * u[None]
    with sx.ScatterDevice("self", dsp) as dev:
        dev.stage(atoms, layout="atoms")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_shells(qv)
    for s in range(len(qlen)):
        r = oracle_mod.compute("self", atoms, b, qv[s], dsp, nthreads=8)
        check_close(fqt[s], fq[s], fq2[s], *r)


@pytest.mark.parametrize("NF", [1, 2, 63, 64, 65, 255, 256, 257, 777, 1500, 2100, 3000, 4096, 5001, 9000])
def test_frame_counts_against_oracle(sx, oracle_mod, NF):
    """Every FFT length class (2^k, 3*2^k, 5*2^k), ragged frame tiles, NF=1."""
    NA = 37
    frames, b = small_trajectory(NA, NF, seed=NF)
    qv = 1.7 * unit_vectors(3, NF)
    with sx.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_shell(qv)
    r = oracle_mod.compute("all", frames, b, qv, nthreads=8)
    check_close(fqt, fq, fq2, *r)


@pytest.mark.parametrize("NF", [1, 2, 63, 64, 65, 257, 1000, 1500, 2000, 2500, 3000, 4000, 5000, 6000])
def test_self_frame_counts_against_oracle(sx, oracle_mod, NF):
    """The self path over the FFT length classes, including series shorter
    than the 64 directly summed tail lags: the half-lengths with a plan fixed
    at compile time (256, 512, 1,024, 1,536, 2,048, 4,096, 5,120) and those
    with the runtime plan (2,560, 3,072, 6,144)."""
    NA = 29
    frames, b = small_trajectory(NA, NF, seed=NF + 1)
    qv = 1.3 * unit_vectors(4, NF + 1)
    atoms = np.ascontiguousarray(frames.transpose(1, 0, 2))
    with sx.ScatterDevice("self") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_shell(qv)
    r = oracle_mod.compute("self", atoms, b, qv, nthreads=8)
    check_close(fqt, fq, fq2, *r)


@pytest.mark.parametrize("vb,split", [(4, 1), (8, 3), (16, 2), (8, 16)])
def test_tile_shapes_agree(sx, oracle_mod, vb, split):
    frames, b = small_trajectory(300, 130, seed=3)
    qv = 0.9 * unit_vectors(37, 4)
    with sx.ScatterDevice("all", vector_block=vb, atom_split=split) as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_shell(qv)
    r = oracle_mod.compute("all", frames, b, qv, nthreads=8)
    check_close(fqt, fq, fq2, *r)


def test_layouts_are_identical(sx):
    frames, b = small_trajectory(70, 90, seed=5)
    qv = 1.3 * unit_vectors(5, 6)
    soa = np.ascontiguousarray(frames.transpose(2, 1, 0))
    atoms = np.ascontiguousarray(frames.transpose(1, 0, 2))
    out = []
    for arr, lay in ((frames, "frames"), (atoms, "atoms"), (soa, "soa")):
        with sx.ScatterDevice("all") as dev:
            dev.stage(arr, layout=lay)
            dev.set_factors(b)
            out.append(dev.compute_shell(qv))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and o[1] == out[0][1] and o[2] == out[0][2]


def test_single_atom_and_single_vector(sx, oracle_mod):
    frames, b = small_trajectory(1, 50, seed=9)
    qv = np.array([[0.3, -0.2, 1.4]])
    for kind in ("all", "self"):
        with sx.ScatterDevice(kind) as dev:
            dev.stage(frames, layout="frames")
            dev.set_factors(b)
            fqt, fq, fq2 = dev.compute_shell(qv)
        coords = frames if kind == "all" else frames.transpose(1, 0, 2)
        check_close(fqt, fq, fq2, *oracle_mod.compute(kind, coords, b, qv))
        assert abs(fqt[0] - b[0] ** 2) < 1e-5 * b[0] ** 2


def test_per_shell_factors(sx, oracle_mod):
    frames, b = small_trajectory(50, 40, seed=11)
    qv = np.array([0.5, 2.0])[:, None, None] * unit_vectors(4, 12)[None]
    bs = np.stack([b, b * np.linspace(0.5, 1.5, b.size)])
    for kind in ("all", "self"):
        with sx.ScatterDevice(kind) as dev:
            dev.stage(frames, layout="frames")
            fqt, fq, fq2 = dev.compute_shells(qv, b_shell=bs)
        coords = frames if kind == "all" else frames.transpose(1, 0, 2)
        for s in range(2):
            check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute(kind, coords, bs[s], qv[s]))


def test_errors_are_reported(sx):
    with sx.ScatterDevice("all") as dev:
        with pytest.raises(sx.SxError) as e:
            dev.compute_shell(np.ones((2, 3)))
        assert e.value.code == -3
        frames, b = small_trajectory(10, 20, seed=1)
        dev.stage(frames, layout="frames")
        with pytest.raises(sx.SxError):
            dev.compute_shell(np.ones((2, 3)))  # no factors yet
        dev.set_factors(b)
        with pytest.raises(sx.SxError) as e:
            dev.compute_shell(np.array([[np.nan, 0, 0]]))
        assert e.value.code == -1
        with pytest.raises(sx.SxError):
            dev.set_factors(b[:5])
    with sx.ScatterDevice("all") as dev:
        frames, b = small_trajectory(2, 163841, seed=1)
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        with pytest.raises(sx.SxError) as e:
            dev.compute_shell(np.ones((1, 3)))
        assert e.value.code == -5


def test_c2_full_size_subset_against_oracle(sx, oracle_mod):
    """Headline case at full size (20000 atoms x 2000 frames): two vectors of
    the largest shell, and a repeat call that must reproduce bit for bit."""
    from sassena_amd import synth
    T = synth.case("c2")
    frames = synth.to_frames_aos(T.xyz)
    qv = T.qvectors(T.nshells - 1)[:2]
    with sx.ScatterDevice("all") as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        r1 = dev.compute_shell(qv)
        r2 = dev.compute_shell(qv)
    assert np.array_equal(r1[0], r2[0])
    ref = oracle_mod.compute("all", frames, T.b, qv, nthreads=8)
    check_close(*r1, *ref)


@pytest.mark.parametrize("NM,ns,split", [(1, 1, 0), (7, 3, 0), (9, 17, 2), (50, 9, 5)])
def test_oriented_ragged_groups(sx, oracle_mod, NM, ns, split):
    """Orientation and shell counts that do not fill the 8 x 8 vector tiles."""
    frames, b = small_trajectory(90, 70, seed=NM + ns)
    u = unit_vectors(NM, ns)
    qlen = np.linspace(0.2, 3.5, ns)
    bs = np.stack([b * (1 + 0.1 * s) for s in range(ns)])
    for per_shell in (False, True):
        with sx.ScatterDevice("all", atom_split=split) as dev:
            dev.stage(frames, layout="frames")
            dev.set_factors(b)
            fqt, fq, fq2 = dev.compute_oriented(qlen, u, b_shell=bs if per_shell else None)
        for s in range(ns):
            bb = bs[s] if per_shell else b
            check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute("all", frames, bb, qlen[s] * u))


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5])
def test_shell_progression_edge_lists(sx, oracle_mod, shape):
    """Linearly spaced lists the stepped-phasor path must handle on every tile
    shape: descending, starting at |q| = 0, two shells, and 21 shells (a
    partial last tile whose middle shell lies past the list), with shared and
    per-shell scattering factors."""
    frames, b = small_trajectory(150, 97, seed=3)
    u = unit_vectors(9, 4)
    lists = [np.linspace(3.0, 0.5, 21), np.linspace(0.0, 2.0, 11), np.array([0.7, 1.9]),
             np.linspace(0.05, 4.0, 21)]
    for qlen in lists:
        ns = len(qlen)
        bs = np.stack([b * (1 + 0.05 * s) for s in range(ns)])
        for per_shell in (False, True):
            with sx.ScatterDevice("all", tile_shape=shape) as dev:
                dev.stage(frames, layout="frames")
                dev.set_factors(b)
                fqt, fq, fq2 = dev.compute_oriented(qlen, u, b_shell=bs if per_shell else None)
                assert dev.timing()["shell_progression"] == 1
            for s in range(ns):
                bb = bs[s] if per_shell else b
                check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute("all", frames, bb, qlen[s] * u))


def test_c2_full_size_oriented_subset(sx, oracle_mod):
    """Full-size headline trajectory, oriented entry point, three shells x two
    orientations, against the oracle."""
    from sassena_amd import synth
    T = synth.case("c2")
    frames = synth.to_frames_aos(T.xyz)
    qlen = T.qlen[[0, 20, 39]]
    u = T.orient[:2]
    with sx.ScatterDevice("all") as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        fqt, fq, fq2 = dev.compute_oriented(qlen, u)
    for s in range(3):
        check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute("all", frames, T.b, qlen[s] * u, nthreads=8))


def test_c2_full_headline_properties(sx):
    """The whole headline workload (20000 atoms x 2000 frames, 40 shells x 200
    orientations) through the oriented entry point, checked by properties that
    hold at any size: a repeat call reproduces bit for bit; tile shapes 1 and 5
    (different tiles, atom ranges and summation orders) agree to the parity
    tolerance; the explicit-vector entry point (a different kernel, one FP64
    phase per vector) agrees on every fifth shell; F(q, 0) is the mean
    |A_v(t)|^2 over vectors and frames, so it is real and positive; fq and fq2
    are the means over vectors of a_v = <C_v(tau)>_tau and of |a_v|^2, so
    fq2 >= |fq|^2."""
    from sassena_amd import synth
    T = synth.case("c2")
    res = {}
    for shape in (5, 1):
        with sx.ScatterDevice("all", tile_shape=shape) as dev:
            dev.stage(T.xyz, layout="soa")
            dev.set_factors(T.b)
            res[shape] = dev.compute_oriented(T.qlen, T.orient)
            if shape == 5:
                again = dev.compute_oriented(T.qlen, T.orient)
                for x, y in zip(res[shape], again):
                    assert np.array_equal(x, y)
                vec = {s: dev.compute_shell(T.qvectors(s)) for s in range(0, T.nshells, 5)}
    fqt, fq, fq2 = res[5]
    assert fqt.shape[0] == T.nshells == 40 and T.orient.shape[0] == 200
    for s in range(T.nshells):
        check_close(fqt[s], fq[s], fq2[s], res[1][0][s], res[1][1][s], res[1][2][s])
        f0 = fqt[s][0]
        assert f0.real > 0 and abs(f0.imag) <= 1e-5 * f0.real
        assert fq2[s].real >= abs(fq[s]) ** 2 * (1 - 1e-5)
    for s, (vfqt, vfq, vfq2) in vec.items():
        check_close(fqt[s], fq[s], fq2[s], vfqt, vfq, vfq2)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5])
def test_shell_progression_recurrence(sx, oracle_mod, shape):
    """Linearly spaced shells take the stepped-phasor path (two sin/cos pairs
    per orientation, atom and frame, the three-term recurrence for the other
    shells of a tile); a perturbed list takes the direct path.  Full-size
    headline trajectory, 40 shells up to 4 A^-1 (up to 19 recurrence steps in
    the widest tile), every tile shape."""
    from sassena_amd import synth
    T = synth.case("c2")
    frames = synth.to_frames_aos(T.xyz)
    u = T.orient[:2]
    ap = np.linspace(0.1, 4.0, 40)
    off = ap.copy()
    off[5] += 1e-3
    refs = {}
    for qlen in (ap, off):
        with sx.ScatterDevice("all", tile_shape=shape) as dev:
            dev.stage(T.xyz, layout="soa")
            dev.set_factors(T.b)
            fqt, fq, fq2 = dev.compute_oriented(qlen, u)
        for s in (0, 5, 9, 15, 19, 31, 39):
            key = float(qlen[s])
            if key not in refs:
                refs[key] = oracle_mod.compute("all", frames, T.b, qlen[s] * u, nthreads=8)
            check_close(fqt[s], fq[s], fq2[s], *refs[key])


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5])
def test_scan_shells_residual_steps(sx, oracle_mod, shape):
    """The headline shells as the reference builds a q scan (interior points
    rounded by single-precision powf, about 6e-8 relative off the exact
    grid): the stepped-phasor path with first-order residual corrections,
    on every tile shape, full-size headline trajectory, two orientations;
    and a list perturbed by 4e-7 relative (residual phases 20x larger) on
    the same path.  Each tile steps on a line through its middle shell; the
    scan's powf points form exact runs (shells 10-19 and 20-39 each lie on one
    arithmetic progression), and the walks from the middle that stay in a run
    need no corrections: with the middle shells, 25, 34, 25, 17 and 30 shells
    so stepped for the shapes 2x8, 4x4, 1x8, 1x16 and 1x20; on the perturbed
    list the middle shells and the 4x4 tiles' one-shell up walks."""
    from sassena_amd import synth
    T = synth.case("c2")
    frames = synth.to_frames_aos(T.xyz)
    u = T.orient[:2]
    scan = synth.linear_shells(0.1, 4.0, 40)
    assert np.max(np.abs(scan - synth.exact_linear_shells(0.1, 4.0, 40))) > 1e-8
    bent = synth.exact_linear_shells(0.1, 4.0, 40) * (1 + 4e-7 * np.cos(np.arange(40)))
    exact_shells = {1: 25, 2: 34, 3: 25, 4: 17, 5: 30}[shape]
    bent_shells = {1: 5, 2: 20, 3: 5, 4: 2, 5: 2}[shape]
    refs = {}
    for qlen in (scan, bent):
        with sx.ScatterDevice("all", tile_shape=shape) as dev:
            dev.stage(T.xyz, layout="soa")
            dev.set_factors(T.b)
            fqt, fq, fq2 = dev.compute_oriented(qlen, u)
            tm = dev.timing()
            assert tm["shell_progression"] == 1 and tm["shell_residuals"] == 1
            assert tm["shell_exact_runs"] == (exact_shells if qlen is scan else bent_shells)
        for s in (1, 5, 9, 13, 17, 20, 27, 33, 38, 39):
            key = float(qlen[s])
            if key not in refs:
                refs[key] = oracle_mod.compute("all", frames, T.b, qlen[s] * u, nthreads=8)
            check_close(fqt[s], fq[s], fq2[s], *refs[key])


def test_residual_bound_falls_back_to_direct(sx, oracle_mod):
    """Residuals whose phase over the coordinates could reach 1e-3 rad take
    the direct path (one phase per shell)."""
    from sassena_amd import synth
    frames, b = small_trajectory(60, 80, seed=21, box=20.0)
    frames = frames + np.float32(40000.0)   # |r| ~ 7e4 A: residual phases near 1e-2 rad
    u = unit_vectors(3, 21)
    qlen = synth.linear_shells(0.1, 2.0, 10)
    with sx.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_oriented(qlen, u)
        assert dev.timing()["shell_progression"] == 0
    for s in (0, 4, 9):
        check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute("all", frames, b, qlen[s] * u))


def test_c5_full_size_subset_against_oracle(sx, oracle_mod):
    """The control-variate workload's trajectory at full size (100,000 atoms
    x 10,000 frames in a 100 A box, pad 20,480) through the oriented entry
    point: the first, a middle and the last of its 60 scan shells x two
    orientations against the FP64 oracle."""
    from sassena_amd import synth
    T = synth.case("c5")
    assert (T.NA, T.NF, T.nshells) == (100000, 10000, 60) and sx.pad_length(T.NF) == 20480
    shells = [0, 29, 59]
    u = T.orient[:2]
    with sx.ScatterDevice("all") as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        fqt, fq, fq2 = dev.compute_oriented(T.qlen[shells], u)
    frames = synth.to_frames_aos(T.xyz)
    del T.xyz
    for i, s in enumerate(shells):
        ref = oracle_mod.compute("all", frames, T.b, T.qlen[s] * u, nthreads=16)
        print(f"c5 shell {s}: error {shell_error(fqt[i], ref[0])[0]:.3e}")
        check_close(fqt[i], fq[i], fq2[i], *ref)


def test_c3_full_size_self_properties(sx, oracle_mod):
    """The incoherent workload at full size (50000 atoms x 5000 frames, two of
    its shells x 100 orientations): F_s(q, 0) = sum_j b_j^2 at every q, since
    each atom's series has modulus b_j; a repeat call reproduces bit for bit;
    the one-after-the-other half layout agrees with the default to the parity
    tolerance; fq and fq2 sum a_jv = <C_jv(tau)>_tau and |a_jv|^2 over the N
    atoms and average over vectors, so fq2 >= |fq|^2 / N (Cauchy-Schwarz); and
    one orientation of the outer shell against the oracle."""
    from sassena_amd import synth
    T = synth.case("c3")
    qlen = T.qlen[[0, T.nshells - 1]]
    b2 = float(np.sum(T.b.astype(np.float64) ** 2))
    with sx.ScatterDevice("self") as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        fqt, fq, fq2 = dev.compute_oriented(qlen, T.orient)
        again = dev.compute_oriented(qlen, T.orient)
        sub = dev.compute_oriented(qlen[1:], T.orient[:1])
    for x, y in zip((fqt, fq, fq2), again):
        assert np.array_equal(x, y)
    with sx.ScatterDevice("self", fft_halves=1) as dev:
        dev.stage(T.xyz, layout="soa")
        dev.set_factors(T.b)
        h1 = dev.compute_oriented(qlen, T.orient)
    for s in range(2):
        assert abs(fqt[s][0] - b2) <= TOL * b2, (fqt[s][0], b2)
        assert fq2[s].real >= abs(fq[s]) ** 2 / np.count_nonzero(T.b) * (1 - 1e-5)
        check_close(fqt[s], fq[s], fq2[s], h1[0][s], h1[1][s], h1[2][s])
    atoms = np.ascontiguousarray(synth.to_frames_aos(T.xyz).transpose(1, 0, 2))
    ref = oracle_mod.compute("self", atoms, T.b, qlen[1] * T.orient[:1], nthreads=16)
    check_close(sub[0][0], sub[1][0], sub[2][0], *ref)


@pytest.mark.parametrize("halves", [3, 2, 1])
@pytest.mark.parametrize("NF", [40, 300, 2000, 5000, 7000, 9000])
def test_self_half_layouts(sx, oracle_mod, NF, halves):
    """The three ways of transforming the two halves of a padded series
    (different blocks, one sweep, one after the other) against the oracle,
    with zero-length atoms in the list and few pairs, so that the last lags
    (averages of a few products, taken from the direct FP64 tail sums) are
    tested at the 1e-5 bound."""
    frames, b = small_trajectory(40, NF, seed=NF)
    b[::3] = 0.0
    u = unit_vectors(6, 7)
    qlen = np.array([0.4, 1.7, 3.1])
    atoms = np.ascontiguousarray(frames.transpose(1, 0, 2))
    with sx.ScatterDevice("self", fft_halves=halves) as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_oriented(qlen, u)
    for s in range(3):
        ref = oracle_mod.compute("self", atoms, b, qlen[s] * u, nthreads=8)
        print(f"NF={NF} halves={halves} shell {s}: error {shell_error(fqt[s], ref[0])[0]:.3e}")
        check_close(fqt[s], fq[s], fq2[s], *ref)


@pytest.mark.parametrize("NF", [10241, 13000, 20000, 31000, 50000])
@pytest.mark.parametrize("kind", ["all", "self"])
def test_long_series_against_oracle(sx, oracle_mod, kind, NF):
    """Frame counts above the largest one-block FFT (rows of N1 = 2, 3, 5
    one-block transforms formed by an on-the-fly column DFT) against the
    FP64 oracle's 2NF transform, coherent and incoherent, with the direct
    method's conjugate on one case."""
    NA = 9 if kind == "all" else 6
    frames, b = small_trajectory(NA, NF, seed=NF + 3)
    if kind == "self":
        b[1] = 0.0
    qv = 1.1 * unit_vectors(3, NF)
    coords = frames if kind == "all" else np.ascontiguousarray(frames.transpose(1, 0, 2))
    methods = ["fftw", "direct"] if NF == 20000 else ["fftw"]
    for method in methods:
        with sx.ScatterDevice(kind, "autocorrelate", method) as dev:
            dev.stage(frames, layout="frames")
            dev.set_factors(b)
            fqt, fq, fq2 = dev.compute_shells(np.stack([qv, 2.3 * qv / 1.1]))
        for s, qq in enumerate((qv, 2.3 * qv / 1.1)):
            ref = oracle_mod.compute(kind, coords, b, qq, "autocorrelate", method, nthreads=8)
            print(f"{kind} NF={NF} {method} shell {s}: error {shell_error(fqt[s], ref[0])[0]:.3e}")
            check_close(fqt[s], fq[s], fq2[s], *ref)


@pytest.mark.parametrize("dsp", ["square", "plain"])
def test_self_series_over_frame_tiles(sx, oracle_mod, dsp):
    """Incoherent square / plain series longer than one LDS tile of frames
    (8,192): the per-pair means are summed over tiles."""
    frames, b = small_trajectory(11, 20000, seed=44)
    atoms = np.ascontiguousarray(frames.transpose(1, 0, 2))
    qv = 0.9 * unit_vectors(4, 44)
    with sx.ScatterDevice("self", dsp) as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_shell(qv)
    check_close(fqt, fq, fq2, *oracle_mod.compute("self", atoms, b, qv, dsp, nthreads=8))


def test_long_series_multipole_and_oriented(sx, oracle_mod):
    """A long series (NF = 12,000) through the oriented entry point (scan
    shells, residual steps) and the sphere multipole."""
    from sassena_amd import synth
    frames, b = small_trajectory(15, 12000, seed=45, box=6.0)
    frames = frames - np.float32(3.0)
    u = unit_vectors(3, 45)
    qlen = synth.linear_shells(0.3, 2.4, 5)
    mom = oracle_mod.sphere_moments(3)
    with sx.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        fqt, fq, fq2 = dev.compute_oriented(qlen, u)
        mfqt, mfq, mfq2 = dev.compute_multipole("sphere", [[0, 0, 1.2]], mom)
    for s in (0, 2, 4):
        check_close(fqt[s], fq[s], fq2[s], *oracle_mod.compute("all", frames, b, qlen[s] * u, nthreads=8))
    check_close(mfqt[0], mfq[0], mfq2[0], *oracle_mod.compute_multipole("sphere", frames, b, [0, 0, 1.2], mom))
