"""Pins the numpy restatement of the control-variate network: its analytic
backward against central finite differences (unrounded FP64), and its BF16
rounding against exact bit patterns."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_bits():
    # bfloat16 keeps 8 significant bits: spacing 2^-7 in [1, 2)
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0],
                 dtype=np.float32)
    r = mlp.bf16(x)
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, 1.0 + 2 ** -7, -2.5, 0.0]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, mlp.NPARAM)
    x = rng.normal(size=(37, mlp.IN))
    t = rng.normal(size=37)
    g, _ = mlp.grads(p, x, t, rounding=False)
    idx = list(rng.choice(mlp.NPARAM, 40, replace=False)) + [mlp.OFF_B3, mlp.OFF_W3, mlp.OFF_B1, mlp.OFF_W1]
    fd = mlp.grads_fd(p, x, t, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_rounded_gradients_close_to_exact():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, mlp.NPARAM)
    x = rng.normal(size=(64, mlp.IN))
    t = rng.normal(size=64)
    g0, _ = mlp.grads(p, x, t, rounding=False)
    g1, _ = mlp.grads(p, x, t, rounding=True)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)
