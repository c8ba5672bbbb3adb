"""Oracle of the learned space adapter network (TEST INFRASTRUCTURE ONLY).

An 8 -> 64 -> 64 -> 1 MLP with tanh-form GELU hidden layers and a logit output, trained
with binary cross-entropy (the network has no reference counterpart; NumCME.jl's adapters
are rule-based). Parameters are one flat vector in the library's order: W1 [8][64], b1,
W2 [64][64], b2, w3 [64], b3.

``forward`` / ``grad`` compute in FP64. With ``bf16=True`` every matrix-multiply operand is
first rounded to BF16 (round to nearest even), at exactly the points where the library's
MFMA kernels round: X, W1, H1, W2, H2 and w3 in the forward pass; H1 and G2 for dW2, G2 and
W2 for G1, X and G1 for dW1 in the backward pass. Without rounding the gradient is pinned by
central differences (``tests/test_oracle_mlp.py``).
"""
from __future__ import annotations

import numpy as np

F, H = 8, 64
NP = F * H + H + H * H + H + H + 1
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    th = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * C_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def forward(p, X, bf16_ops=False, keep=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    Z1 = r(X) @ r(W1) + b1
    H1 = gelu(Z1)
    Z2 = r(H1) @ r(W2) + b2
    H2 = gelu(Z2)
    z3 = r(H2) @ r(w3) + b3
    if keep:
        return z3, (X, Z1, H1, Z2, H2)
    return z3


def loss(p, X, Y, bf16_ops=False):
    z = forward(p, X, bf16_ops)
    Y = np.asarray(Y, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0) - z * Y + np.log1p(np.exp(-np.abs(z)))))


def grad(p, X, Y, bf16_ops=False):
    """Gradient of the mean binary cross-entropy (flat, library order) and the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    z3, (X, Z1, H1, Z2, H2) = forward(p, X, bf16_ops, keep=True)
    Y = np.asarray(Y, dtype=np.float64)
    n = X.shape[0]
    d3 = 1.0 / (1.0 + np.exp(-z3)) - Y
    gw3 = H2.T @ d3
    gb3 = d3.sum()
    G2 = d3[:, None] * w3[None, :] * gelu_d(Z2)
    gW2 = r(H1).T @ r(G2)
    gb2 = G2.sum(axis=0)
    G1 = (r(G2) @ r(W2).T) * gelu_d(Z1)
    gW1 = r(X).T @ r(G1)
    gb1 = G1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]]) / n
    L = float(np.mean(np.maximum(z3, 0) - z3 * Y + np.log1p(np.exp(-np.abs(z3)))))
    return g, L
