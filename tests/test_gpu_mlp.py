"""The GPU forward/backward of the control-variate network against the FP64
oracle (same BF16 roundings), and training on a smooth target."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows", [16, 64, 100, 1000])
def test_gradients_against_oracle(rows):
    from sassena_amd.mlp import MLP
    rng = np.random.default_rng(rows)
    x = rng.normal(size=(rows, 16)).astype(np.float32)
    t = rng.normal(size=rows).astype(np.float32)
    with MLP(seed=3) as net:
        p = net.params()
        g, loss = net.gradients(x, t)
    rg, rloss = omlp.grads(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - rloss) <= 1e-4 * rloss
    # per block of parameters: relative error of the gradient vector
    for a, b in ((omlp.OFF_W1, omlp.OFF_B1), (omlp.OFF_B1, omlp.OFF_W2), (omlp.OFF_W2, omlp.OFF_B2),
                 (omlp.OFF_B2, omlp.OFF_W3), (omlp.OFF_W3, omlp.OFF_B3), (omlp.OFF_B3, omlp.NPARAM)):
        err = np.linalg.norm(g[a:b] - rg[a:b]) / max(np.linalg.norm(rg[a:b]), 1e-30)
        assert err < 2e-3, (a, b, err)


def test_predict_against_oracle():
    from sassena_amd.mlp import MLP
    rng = np.random.default_rng(5)
    x = rng.normal(size=(333, 16)).astype(np.float32)
    with MLP(seed=9) as net:
        y = net.predict(x)
        p = net.params()
    ry, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(y - ry)) < 1e-4 * max(1.0, np.max(np.abs(ry)))


def test_training_fits_a_smooth_function():
    from sassena_amd.mlp import MLP
    rng = np.random.default_rng(7)
    x = np.zeros((20000, 16), np.float32)
    x[:, :2] = rng.uniform(-1, 1, size=(20000, 2))
    t = (np.exp(-x[:, 0] ** 2 * 2) * np.cos(2 * x[:, 1])).astype(np.float32)
    with MLP(seed=1) as net:
        y0 = net.predict(x)
        net.train(x, t, steps=1500, batch=2048, lr=3e-3)
        y1 = net.predict(x)
    before = np.mean((y0 - t) ** 2)
    after = np.mean((y1 - t) ** 2)
    assert after < 1e-3 and after < 0.01 * before
