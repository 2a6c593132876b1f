"""The allocation-free entry point (sx_compute_shells_oriented_noalloc)
against the allocating one (sx_compute_shells_oriented) through the C ABI:
two calls in a row on one context, the second with fewer shells and
orientations, each within the parity bound of the other and of the FP64
oracle; a call whose single shell exceeds the scratch maximum fails cleanly."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from util import fq2_bound, shell_error, small_trajectory, unit_vectors  # noqa: E402

pytestmark = pytest.mark.gpu


def test_noalloc_matches_the_allocating_entry_point(oracle_mod):
    import sassena_amd
    from sassena_amd import synth
    frames, b = small_trajectory(700, 900, seed=31, box=25.0)
    u = unit_vectors(40, 31)
    calls = [(synth.linear_shells(0.2, 3.5, 24), u), (np.array([0.7, 1.3, 2.9]), u[:5])]
    with sassena_amd.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        for qlen, uu in calls:
            ref = dev.compute_oriented(qlen, uu)
            got = dev.compute_oriented(qlen, uu, noalloc=True)
            assert got[0].shape == ref[0].shape
            err = shell_error(got[0], ref[0])
            assert np.max(err) < 1e-5, err
            scale = np.abs(ref[0][:, 0])
            assert np.all(np.abs(got[1] - ref[1]) <= 1e-5 * scale)
            assert all(abs(g - r) <= fq2_bound(r, f0) for g, r, f0 in zip(got[2], ref[2], scale))
            s = len(qlen) - 1
            o = oracle_mod.compute("all", frames, b, qlen[s] * uu, nthreads=8)
            assert shell_error(got[0][s], o[0])[0] < 1e-5


def test_noalloc_rejects_a_shell_above_the_scratch():
    import sassena_amd
    frames, b = small_trajectory(8, 2000, seed=32)
    u = unit_vectors(20000, 32)
    with sassena_amd.ScatterDevice("all") as dev:
        dev.stage(frames, layout="frames")
        dev.set_factors(b)
        with pytest.raises(sassena_amd.SxError) as e:
            dev.compute_oriented([1.0], u, noalloc=True)
        assert e.value.code == -4 and "scratch" in str(e.value)
        # the context still works afterwards, both ways
        r1 = dev.compute_oriented([1.0, 2.0], u[:3])
        r2 = dev.compute_oriented([1.0, 2.0], u[:3], noalloc=True)
        assert np.max(shell_error(r2[0], r1[0])) < 1e-5
