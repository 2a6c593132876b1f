"""The batched multi-trajectory path (one captured graph, persistent
amplitude kernel): every trajectory against the FP64 oracle and against the
single-trajectory entry point; each trajectory alone through its own graph."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import shell_error, small_trajectory, unit_vectors  # noqa: E402

pytestmark = pytest.mark.gpu


def uneven(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        NA = int(rng.integers(20, 400))
        NF = int(rng.integers(40, 700))
        frames, b = small_trajectory(NA, NF, seed=seed * 100 + k)
        ns = int(rng.integers(1, 12))
        NM = int(rng.integers(1, 24))
        out.append((frames, b, np.sort(rng.uniform(0.1, 4.0, ns)), unit_vectors(NM, k + seed)))
    return out


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_batch_against_oracle(oracle_mod, shape):
    import sassena_amd
    trajs = uneven(6, 17)
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
            assert abs(fq2[s] - rfq2) < 1e-5 * (abs(rfq2) + 1e-3 * abs(rfqt[0]) ** 2)


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
