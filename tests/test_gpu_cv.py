"""The learned control variate end to end on a scaled-down large case:
trained on the first 20 % of frames, it must move the 16-orientation
estimate towards the plain many-orientation average on the remaining
frames, without a systematic bias."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_control_variate_gains_over_small_estimate():
    from sassena_amd import cv, synth
    rng = np.random.default_rng(55)
    NA, NF = 3000, 1200                        # density of the large case (0.1 atoms / A^3)
    box = (NA / 0.1) ** (1.0 / 3.0)
    b = synth.scattering_lengths(rng, NA)
    xyz = synth.random_walk(rng, NA, NF, box, 0.1)
    orient = synth.orientations(rng, 200)
    qlen = np.linspace(0.3, 5.0, 16)
    r = cv.run(xyz, b, qlen, orient, ntrain_frames=NF // 5, steps=1500, batch=4096)
    print({k: r[k] for k in ("error_small", "error_cv", "bias_cv", "train_loss")})
    assert r["error_cv"]["shell_mean"] < r["error_small"]["shell_mean"]
    assert abs(r["bias_cv"]) < r["error_small"]["shell_mean"]
