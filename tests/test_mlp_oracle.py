"""Pins the float64 predictor oracle (oracle/mlp_ref.py) that the GPU MLP
kernels are checked against: bf16 rounding known answers, analytic gradients
against central finite differences, and the loss against a direct BCE."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40, 0.1], np.float32)
    y = mlp_ref.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.015625 and y[3] == 1.0078125
    assert y[4] == -2.5 and y[7] == np.float64(0.10009765625)
    assert abs(y[5] / 3.0e38 - 1) < 2 ** -8


def setup(nd=5, n=64, seed=1):
    rng = np.random.default_rng(seed)
    P = 19585 + 16 * nd
    params = rng.normal(0, 0.2, P)
    s = np.zeros((n, 8))
    s[:, 0:3] = rng.uniform(0, 4, (n, 3))
    d = rng.normal(size=(n, 3))
    s[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 6] = rng.integers(0, nd, n)
    s[:, 7] = rng.integers(0, 2, n)
    center = np.full(3, 2.0)
    inv = np.full(3, 0.5)
    return params, s, nd, center, inv


def test_analytic_gradient_matches_finite_differences():
    params, s, nd, center, inv = setup()
    g = mlp_ref.analytic_grad_f64(params, s, nd, center, inv)
    rng = np.random.default_rng(3)
    # a few parameters from every block: W1, b1, W2, b2, w3, b3, embeddings
    idx = np.concatenate([rng.integers(0, 2816, 6), rng.integers(2816, 2944, 3),
                          rng.integers(2944, 19328, 6), rng.integers(19328, 19456, 3),
                          rng.integers(19456, 19584, 3), [19584],
                          19585 + 16 * s[:4, 6].astype(int) + rng.integers(0, 16, 4)])
    # small steps keep every ReLU on one side of its kink
    num = mlp_ref.numerical_grad(params, s, nd, center, inv, idx, eps=1e-6)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-8)


def test_bf16_forward_backward_close_to_float64():
    params, s, nd, center, inv = setup(n=128)
    loss, g, z = mlp_ref.forward_backward(params, s, nd, center, inv)
    g64 = mlp_ref.analytic_grad_f64(params, s, nd, center, inv)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 0.05
    # the summed loss is the BCE of the logits
    y = s[:, 7]
    bce = np.logaddexp(0, z) - y * z
    assert abs(loss - bce.sum()) < 1e-6 * abs(loss)
