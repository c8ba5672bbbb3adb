"""GPU parity of the learned adapter's network: the BF16 MFMA forward and backward kernels
against the oracle (which rounds the same operands to BF16 and computes in FP64), against the
unrounded FP64 network, and a training run on a separable rule.
"""
import numpy as np
import pytest

import numcme_hip as nh
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, om.F)).astype(np.float32)
    Y = (rng.random(n) < 0.3).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000])
def test_forward_matches_oracle(gpu, n):
    net = nh.AdapterMlp(seed=3)
    X, _ = _data(n, n)
    z = net.forward(X).astype(np.float64)
    p = net.params.astype(np.float64)
    zo = om.forward(p, X, bf16_ops=True)
    # same BF16 operands; the kernel keeps hidden activations in FP32 before rounding them,
    # so an activation can land on the other side of a BF16 rounding boundary now and then
    assert np.max(np.abs(z - zo)) <= 2e-3 * (1 + np.max(np.abs(zo)))
    assert np.median(np.abs(z - zo)) <= 1e-5 * (1 + np.max(np.abs(zo)))
    # plain FP64 network: the BF16 operand error bound
    zf = om.forward(p, X, bf16_ops=False)
    assert np.max(np.abs(z - zf)) <= 5e-2 * (1 + np.max(np.abs(zf)))


def test_forward_empty(gpu):
    net = nh.AdapterMlp(seed=1)
    assert net.forward(np.zeros((0, om.F))).shape == (0,)


@pytest.mark.parametrize("n", [1, 65, 777, 4096])
def test_gradient_matches_oracle(gpu, n):
    net = nh.AdapterMlp(seed=5)
    rng = np.random.default_rng(7)
    net.params = (net.params + rng.normal(scale=0.05, size=om.NP)).astype(np.float32)  # nonzero biases
    X, Y = _data(n, 100 + n)
    g, loss = net.grad(X, Y)
    p = net.params.astype(np.float64)
    go, lo = om.grad(p, X, Y, bf16_ops=True)
    assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
    scale = np.max(np.abs(go))
    assert np.max(np.abs(g - go)) <= 2e-3 * scale
    assert np.linalg.norm(g - go) <= 2e-3 * np.linalg.norm(go)
    # unrounded FP64 gradient (pinned by central differences on the CPU)
    gf, _ = om.grad(p, X, Y, bf16_ops=False)
    assert np.linalg.norm(g - gf) <= 3e-2 * np.linalg.norm(gf)


def test_gradient_is_deterministic(gpu):
    net = nh.AdapterMlp(seed=2)
    X, Y = _data(3000, 9)
    g1, l1 = net.grad(X, Y)
    g2, l2 = net.grad(X, Y)
    assert np.array_equal(g1, g2) and l1 == l2


def test_training_learns_a_separable_rule(gpu):
    rng = np.random.default_rng(21)
    X = rng.normal(size=(20000, om.F)).astype(np.float32)
    Y = (X[:, 0] + 0.5 * X[:, 3] - X[:, 5] > 0.2).astype(np.float32)
    net = nh.AdapterMlp(seed=4)
    _, l0 = net.grad(X, Y)
    loss = net.train(X, Y, epochs=15, batch=1024, lr=3e-3, seed=1)
    assert loss < 0.5 * l0
    Xt = rng.normal(size=(5000, om.F)).astype(np.float32)
    Yt = Xt[:, 0] + 0.5 * Xt[:, 3] - Xt[:, 5] > 0.2
    acc = np.mean((net.forward(Xt) > 0) == Yt)
    assert acc > 0.95
    # the device loss agrees with the oracle's at the trained weights
    _, lt = net.grad(X, Y)
    assert abs(lt - om.loss(net.params.astype(np.float64), X, Y, bf16_ops=True)) <= 1e-4


def test_rejects_bad_shapes(gpu):
    net = nh.AdapterMlp()
    with pytest.raises(nh.ArgumentError):
        net.forward(np.zeros((4, 7)))
    with pytest.raises(nh.ArgumentError):
        net.params = np.zeros(10)
