"""Pins the CPU oracle of the round-off MLP -- CPU only: backpropagation equals central
finite differences of the loss, and the forward pass equals an explicit per-sample loop."""
import numpy as np

from mersenneprime_python_cpp_gmp_benchmark_amd import nselect
from oracle import mlp_oracle as M


def test_param_count_and_layout():
    assert M.N_PARAMS == 4481 == len(nselect.init_params(0))


def test_backprop_equals_finite_differences():
    rng = np.random.default_rng(1)
    P = nselect.init_params(3).astype(np.float64) + rng.normal(0, 0.05, M.N_PARAMS)
    X = rng.normal(0, 1, (40, 3))
    t = rng.normal(0, 1, 40)
    g, _ = M.grad(P, X, t)
    idx = list(range(0, 192, 17)) + list(range(192, 256, 7)) + list(range(256, 4352, 331)) + \
        list(range(4352, 4416, 9)) + list(range(4416, 4481, 11)) + [4480]
    fd = M.finite_difference_grad(P, X, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_forward_matches_explicit_loop():
    rng = np.random.default_rng(2)
    P = nselect.init_params(5).astype(np.float64)
    X = rng.normal(0, 1, (5, 3))
    W1, b1, W2, b2, w3, b3 = M.unpack(P)
    for i in range(5):
        h1 = [np.tanh(sum(W1[a, k] * X[i, k] for k in range(3)) + b1[a]) for a in range(64)]
        h2 = [np.tanh(sum(W2[a, b] * h1[b] for b in range(64)) + b2[a]) for a in range(64)]
        y = sum(w3[a] * h2[a] for a in range(64)) + b3
        assert abs(M.forward(P, X[i:i + 1])[0][0] - y) < 1e-12


def test_conservative_length_is_reference_rule():
    from oracle import oracle as O
    for p in (1048573, 16777213, 20000003, 33554393, 2**20 + 7):
        assert nselect.conservative_length(p) == O.fft_length(p, 18.0)


def test_features_and_targets_roundtrip():
    e = np.array([1e-4, 3e-3, 0.2])
    np.testing.assert_allclose(nselect.err_of(nselect.target_of(e)), e, rtol=1e-6)
    f = nselect.features([2**22], [2**19])
    assert f.shape == (1, 3)
