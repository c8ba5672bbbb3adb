"""Pins of the MLP oracle (CPU): back-propagation against central differences, the BF16
rounding helper against exact values."""
import numpy as np

from oracle import mlp


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=mlp.NP)
    return p


def test_bf16_rounding():
    # BF16 keeps 7 fraction bits: near 1 the spacing is 2^-7. 1 + 2^-8 is a tie and goes to the
    # even neighbour 1; 1 + 3*2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even); 1 + 3*2^-9
    # is not a tie and rounds up to 1 + 2^-7.
    x = np.array([1.0, 1 + 2.0**-8, 1 + 3 * 2.0**-8, 1 + 3 * 2.0**-9, -2.5, 3.0e-3], dtype=np.float32)
    np.testing.assert_array_equal(mlp.bf16(x)[:4], [1.0, 1.0, 1 + 2.0**-6, 1 + 2.0**-7])
    assert mlp.bf16(x)[4] == -2.5
    assert abs(mlp.bf16(x)[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_backprop_matches_central_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40, mlp.F))
    Y = (rng.random(40) < 0.4).astype(float)
    p = _params()
    g, _ = mlp.grad(p, X, Y)
    idx = rng.choice(mlp.NP, size=120, replace=False)
    idx = np.concatenate([idx, [mlp.NP - 1, 0, mlp.F * mlp.H]])  # b3, W1[0,0], b1[0]
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss(pp, X, Y) - mlp.loss(pm, X, Y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), i


def test_gelu_derivative():
    x = np.linspace(-6, 6, 101)
    fd = (mlp.gelu(x + 1e-6) - mlp.gelu(x - 1e-6)) / 2e-6
    np.testing.assert_allclose(mlp.gelu_d(x), fd, atol=1e-8)
