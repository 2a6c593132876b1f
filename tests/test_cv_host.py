"""Host-side pieces of the learned control variate (no GPU): feature rows,
lag clamping, shell errors, and the crystal generator."""
import numpy as np

from sassena_amd import cv, synth


def test_features_layout_and_clamping():
    rng = np.random.default_rng(0)
    ns, NF, tmax = 3, 10, 4.0
    f16 = rng.normal(size=(ns, NF)) + 1j * rng.normal(size=(ns, NF))
    dprev = rng.normal(size=(ns, int(tmax))) + 1j * rng.normal(size=(ns, int(tmax)))
    scale = np.array([2.0, 4.0, 8.0])
    x = cv.features([0.5, 1.0, 2.0], f16, dprev, scale, 1.0, tmax, 2.0)
    assert x.shape == (ns * NF, cv.NFEAT) and x.dtype == np.float32
    r = x.reshape(ns, NF, cv.NFEAT)
    # lags beyond the training range see the last trained lag
    assert np.all(r[:, 3:, 1] == r[:, 3:4, 1]) and np.all(r[:, 3:, 6] == r[:, 3:4, 6])
    assert np.allclose(r[1, :, 3], f16[1].real / 4.0, rtol=1e-6)
    assert np.allclose(r[2, 2, 6], dprev[2, 2].real / 8.0, rtol=1e-6)
    assert np.allclose(r[:, 0, 0], [0.25, 0.5, 1.0])
    assert np.all(x[:, 10:] == 0)


def test_shell_errors():
    ref = np.array([[2.0, 1.0, 0.5], [4.0, 4.0, 4.0]], dtype=complex)
    f = ref + np.array([[0.2, 0.0, -0.1], [0.0, 0.4, 0.0]])
    emax, emean = cv.shell_errors(f, ref)
    assert np.allclose(emax, [0.1, 0.1]) and np.allclose(emean, [0.05, 0.1 / 3])


def test_lattice_vibrations_statistics():
    rng = np.random.default_rng(1)
    xyz = synth.lattice_vibrations(rng, 27, 4000, 2.0, 0.1, 50.0)
    sites = np.round(xyz.mean(axis=2) / 2.0) * 2.0
    disp = xyz - sites[:, :, None]
    assert abs(disp.std() - 0.1) < 0.01            # stationary std
    lag = disp[:, :, 50:] * disp[:, :, :-50]
    assert abs(lag.mean() / disp.var() - np.exp(-1.0)) < 0.1   # correlation time 50 frames
