"""The batched multi-trajectory path (one captured graph, persistent
amplitude kernel): every trajectory against the FP64 oracle and against the
single-trajectory entry point; each trajectory alone through its own graph."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import fq2_bound, shell_error, small_trajectory, unit_vectors  # noqa: E402

pytestmark = pytest.mark.gpu


def uneven(n, seed, long_frames=()):
    """n small trajectories of random size, then one of few atoms, shells and
    orientations for each frame count in long_frames."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        NA = int(rng.integers(20, 400))
        NF = int(rng.integers(40, 700))
        frames, b = small_trajectory(NA, NF, seed=seed * 100 + k)
        ns = int(rng.integers(1, 12))
        NM = int(rng.integers(1, 24))
        out.append((frames, b, np.sort(rng.uniform(0.1, 4.0, ns)), unit_vectors(NM, k + seed)))
    for k, NF in enumerate(long_frames):
        NA = int(rng.integers(20, 60))
        frames, b = small_trajectory(NA, NF, seed=seed * 1000 + k)
        ns = int(rng.integers(1, 3))
        out.append((frames, b, np.sort(rng.uniform(0.1, 4.0, ns)), unit_vectors(2, k + seed)))
    return out


# One frame count per FFT plan above Lh = 1024: half lengths 1536 ... 10240,
# with 8, 16, 20 and 24 points per thread and 64 to 512 threads, up to the
# largest single-block plan (160 KB of LDS).
LONG_FRAMES = (1100, 1800, 2100, 2600, 3500, 5001, 6000, 8000, 9000)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_batch_against_oracle(oracle_mod, shape):
    """Six small trajectories and one per long FFT plan class against the
    FP64 oracle, every tile shape (0: mixed tiles)."""
    import sassena_amd
    assert [sassena_amd.pad_length(nf) // 2 for nf in LONG_FRAMES] == [1536, 2048, 2560, 3072, 4096, 5120, 6144,
                                                                          8192, 10240]
    trajs = uneven(6, 17, LONG_FRAMES if shape in (0, 5) else LONG_FRAMES[2:4])
    with sassena_amd.Batch(tile_shape=shape) as bt:
        bt.stage([(np.ascontiguousarray(f.transpose(2, 1, 0)), b, ql, u) for f, b, ql, u in trajs])
        res = bt.run()
        t = bt.timing()
    assert t["terms"] == sum(len(ql) * u.shape[0] * f.shape[1] * f.shape[0] for f, b, ql, u in trajs)
    for (frames, b, ql, u), (fqt, fq, fq2) in zip(trajs, res):
        for s in range(len(ql)):
            rfqt, rfq, rfq2 = oracle_mod.compute("all", frames, b, ql[s] * u, nthreads=8)
            assert shell_error(fqt[s], rfqt)[0] < 1e-5
            assert abs(fq[s] - rfq) < 1e-5 * abs(rfqt[0])
            assert abs(fq2[s] - rfq2) <= fq2_bound(rfq2, rfqt[0])


@pytest.mark.parametrize("kind", ["random", "exact", "scan"])
def test_mixed_tiles_against_oracle(oracle_mod, kind):
    """Shell counts whose cheapest cover mixes 1x8, 1x16 and 1x20 tiles with a
    partial last tile (7: 8; 10: 16; 21 and 23: 16 + 8; 35: 20 + 16; 44:
    20 + 16 + 8; 40: 20 + 20), the shells at random |q| (one phase each), on
    an exact grid (stepped phasors) or from a single-precision scan (stepped
    with residual corrections; no count has n - 1 a power of two, where the
    scan is exact; the 40-point scan's shells 20-39 form one exact run, whose
    tile is stepped without corrections):
    every shell against the FP64 oracle, and the whole result against fixed
    1 x 20 tiles (to the oracle's bound: the two step phasors from different
    middle shells, and their rounding differs at 1e-6)."""
    import sassena_amd
    from sassena_amd import synth
    rng = np.random.default_rng(71)
    trajs = []
    for k, ns in enumerate((7, 10, 21, 23, 35, 40, 44)):
        frames, b = small_trajectory(int(rng.integers(60, 200)), int(rng.integers(100, 400)), seed=700 + k)
        if kind == "random":
            ql = np.sort(rng.uniform(0.1, 4.0, ns))
        elif kind == "exact":
            ql = synth.exact_linear_shells(0.1, 4.0, ns)
        else:
            ql = synth.linear_shells(0.1, 4.0, ns)
        trajs.append((frames, b, ql, unit_vectors(int(rng.integers(2, 7)), 70 + k)))
    staged = [(np.ascontiguousarray(f.transpose(2, 1, 0)), b, ql, u) for f, b, ql, u in trajs]
    with sassena_amd.Batch() as bt:
        bt.stage(staged)
        res = bt.run()
        tmb = bt.timing()
        assert tmb["tile_shape"] == 5
        assert (tmb["terms_exact_runs"] > 0) == (kind == "scan")
    with sassena_amd.Batch(tile_shape=5) as bt:
        bt.stage(staged)
        fixed = bt.run()
    for (frames, b, ql, u), (fqt, fq, fq2), (xfqt, _, _) in zip(trajs, res, fixed):
        assert np.max(shell_error(fqt, xfqt)) < 1e-5
        for s in range(len(ql)):
            rfqt, rfq, rfq2 = oracle_mod.compute("all", frames, b, ql[s] * u, nthreads=8)
            assert shell_error(fqt[s], rfqt)[0] < 1e-5, (len(ql), s)
            assert abs(fq[s] - rfq) < 1e-5 * abs(rfqt[0])
            assert abs(fq2[s] - rfq2) <= fq2_bound(rfq2, rfqt[0])


def test_batch_matches_single_and_run_one():
    import sassena_amd
    trajs = uneven(5, 23)
    with sassena_amd.Batch() as bt:
        bt.stage([(np.ascontiguousarray(f.transpose(2, 1, 0)), b, ql, u) for f, b, ql, u in trajs])
        res = bt.run()
        again = bt.run()
        lat = []
        for k in range(len(trajs)):
            one, ms = bt.run_one(k)
            lat.append(ms)
            assert np.array_equal(one[0], res[k][0])
    assert all(np.array_equal(a[0], b[0]) for a, b in zip(res, again))
    assert all(ms > 0 for ms in lat)
    for (frames, b, ql, u), (fqt, fq, fq2) in zip(trajs, res):
        with sassena_amd.ScatterDevice("all") as dev:
            dev.stage(frames, layout="frames")
            dev.set_factors(b)
            sfqt, sfq, sfq2 = dev.compute_oriented(ql, u)
        assert np.max(shell_error(fqt, sfqt)) < 1e-6


def test_batch_errors():
    import sassena_amd
    with sassena_amd.Batch() as bt:
        with pytest.raises(sassena_amd.SxError):
            bt.run()
        frames, b = small_trajectory(10, 20000, seed=1)
        with pytest.raises(sassena_amd.SxError) as e:
            bt.stage([(np.ascontiguousarray(frames.transpose(2, 1, 0)), b, [1.0], unit_vectors(2, 1))])
        assert e.value.code == -5


def test_c4_batch_at_full_frame_counts(oracle_mod):
    """The 64-trajectory batch of the uneven workload with its frame counts
    (256-4,096: every FFT plan class the workload reaches), shells
    (10-60, a reference q scan, so stepped phasors, residual-corrected unless
    the scan is exact in single precision) and orientations
    (50-300), atom counts scaled by 0.1: every trajectory of the batch graph
    equals its own graph (sx_batch_run_one) bit for bit and the
    single-trajectory entry point to 1e-6; the trajectories of the longest
    frame counts match the FP64 oracle on their first, middle and last
    shells (two orientations)."""
    from sassena_amd import Batch, ScatterDevice, synth
    trajs = synth.c4_batch(scale=0.1, frame_scale=1.0)
    assert len(trajs) == 64
    bt = Batch()
    bt.stage([(T.xyz, T.b, T.qlen, T.orient) for T in trajs])
    res = bt.run()
    t = bt.timing()
    assert t["terms"] == sum(T.terms() for T in trajs)
    for k, T in enumerate(trajs):
        one, _ = bt.run_one(k)
        assert np.array_equal(one[0], res[k][0]) and np.array_equal(one[1], res[k][1])
        with ScatterDevice("all") as dev:
            dev.stage(T.xyz, layout="soa")
            dev.set_factors(T.b)
            sfqt, sfq, sfq2 = dev.compute_oriented(T.qlen, T.orient)
            # scans with n - 1 a power of two are exact; the others carry residuals
            assert dev.timing()["shell_progression"] == 1
        assert np.max(shell_error(res[k][0], sfqt)) < 1e-6, k
    bt.close()
    longest = sorted(range(64), key=lambda k: trajs[k].NF)[-3:]
    for k in longest:
        T = trajs[k]
        frames = synth.to_frames_aos(T.xyz)
        u = T.orient[:2]
        with ScatterDevice("all") as dev:
            dev.stage(T.xyz, layout="soa")
            dev.set_factors(T.b)
            fqt, fq, fq2 = dev.compute_oriented(T.qlen, u)
        for s in (0, T.nshells // 2, T.nshells - 1):
            rfqt, rfq, rfq2 = oracle_mod.compute("all", frames, T.b, T.qlen[s] * u, nthreads=8)
            assert shell_error(fqt[s], rfqt)[0] < 1e-5, (k, s)
            assert abs(fq[s] - rfq) < 1e-5 * abs(rfqt[0])


def test_time_amplitude_leaves_results():
    """The amplitude-only timing call does not disturb the next batch run."""
    from sassena_amd import Batch, synth
    trajs = synth.c4_batch(count=3, scale=0.05)
    bt = Batch()
    bt.stage([(T.xyz, T.b, T.qlen, T.orient) for T in trajs])
    first = bt.run()
    ms = bt.time_amplitude()
    assert ms > 0
    again = bt.run()
    for a, b in zip(first, again):
        assert np.array_equal(a[0], b[0])
    bt.close()
