"""Pins the oracle MLP (oracle/mlp.py): analytic gradients against central
finite differences, GELU against its definition, bf16 rounding against known
bit patterns."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    nin, nh = 5, 7
    rng = np.random.default_rng(3)
    th = rng.standard_normal(mlp.nparams(nin, nh)) * 0.5
    x = rng.standard_normal((9, nin))
    t = rng.standard_normal(9)
    _, g = mlp.loss_and_grad(th, x, t, nin, nh)
    eps = 1e-6
    num = np.zeros_like(th)
    for i in range(th.size):
        e = np.zeros_like(th)
        e[i] = eps
        num[i] = (mlp.loss_and_grad(th + e, x, t, nin, nh)[0] - mlp.loss_and_grad(th - e, x, t, nin, nh)[0]) / (2 * eps)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def test_gelu_definition():
    z = np.linspace(-4, 4, 17)
    from scipy.stats import norm
    np.testing.assert_allclose(mlp.gelu(z), z * norm.cdf(z), rtol=1e-12, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(mlp.gelu_grad(z), (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_bf16_rounding():
    assert mlp.bf16(1.0) == 1.0
    assert mlp.bf16(1.0 + 2 ** -8) == 1.0                     # tie -> even
    assert mlp.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7       # above half -> up
    assert mlp.bf16(-2.0 ** 100) == -2.0 ** 100
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.max(np.abs(mlp.bf16(x) - x) / np.abs(x)) <= 2 ** -8
