"""CPU oracle of the round-off MLP (3 -> 64 -> 64 -> 1, tanh), float64 numpy.

TEST INFRASTRUCTURE ONLY: the GPU kernels' forward pass, gradients and training are checked
against this restatement; it is itself pinned by central finite differences
(tests/test_mlp_oracle.py).  Parameter layout matches the kernels: W1[64][3], b1[64],
W2[64][64], b2[64], w3[64], b3.
"""
from __future__ import annotations

import numpy as np

H = 64
N_PARAMS = H * 3 + H + H * H + H + H + 1


def unpack(P):
    P = np.asarray(P, np.float64)
    o = 0
    W1 = P[o:o + 3 * H].reshape(H, 3); o += 3 * H
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].reshape(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    w3 = P[o:o + H]; o += H
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def forward(P, X):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, (h1, h2)


def loss(P, X, t):
    y, _ = forward(P, X)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grad(P, X, t):
    """(gradient, loss) of the mean squared error, by backpropagation."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    t = np.asarray(t, np.float64)
    n = len(t)
    y, (h1, h2) = forward(P, X)
    dy = 2.0 * (y - t) / n
    dz2 = np.outer(dy, w3) * (1.0 - h2 ** 2)
    dz1 = (dz2 @ W2) * (1.0 - h1 ** 2)
    g = np.concatenate([(dz1.T @ X).ravel(), dz1.sum(0), (dz2.T @ h1).ravel(), dz2.sum(0), dy @ h2, [dy.sum()]])
    return g, float(np.mean((y - t) ** 2))


def finite_difference_grad(P, X, t, idx, h=1e-6):
    P = np.asarray(P, np.float64).copy()
    out = []
    for i in idx:
        a = P.copy(); a[i] += h
        b = P.copy(); b[i] -= h
        out.append((loss(a, X, t) - loss(b, X, t)) / (2 * h))
    return np.array(out)
