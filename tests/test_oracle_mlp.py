"""Pins the CPU restatement of the signed-distance surrogate (oracle/mlp.py):
bf16 rounding against hand-computed ties, and the backward pass of the exact
float64 network against central finite differences."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp  # noqa: E402


def test_bf16_round_to_nearest_even():
    e = 2.0 ** -7  # bf16 spacing in [1, 2)
    cases = {
        1.0: 1.0,
        1.0 + e / 2: 1.0,              # tie -> even (mantissa 0)
        1.0 + e + e / 2: 1.0 + 2 * e,  # tie -> even (mantissa 2)
        1.0 + e / 2 + 2.0 ** -20: 1.0 + e,  # just above the tie
        -3.0: -3.0,
        0.0: 0.0,
        2.0 ** -130: 2.0 ** -130,      # subnormal, exact in bf16
    }
    for x, want in cases.items():
        assert mlp.bf16(np.array([x], np.float32))[0] == np.float32(want), x
    assert np.isnan(mlp.bf16(np.array([np.nan], np.float32))[0])
    assert np.isinf(mlp.bf16(np.array([np.inf], np.float32))[0])


def test_layout():
    assert mlp.NUM_PARAMS == 256 * 32 + 256 + 3 * (256 * 256 + 256) + 256 + 1
    p = np.arange(mlp.NUM_PARAMS, dtype=np.float64)
    np.testing.assert_array_equal(mlp.pack(mlp.unpack(p)), p)


def _random_params(rng, in_dim):
    d = {}
    for l in range(1, 5):
        k = 32 if l == 1 else 256
        fan = in_dim if l == 1 else 256
        W = rng.uniform(-1, 1, (256, k)) / np.sqrt(fan)
        if l == 1:
            W[:, in_dim:] = 0.0
        d[f"W{l}"] = W
        d[f"b{l}"] = rng.uniform(-1, 1, 256) / np.sqrt(fan)
    d["w5"] = rng.uniform(-1, 1, 256) / 16.0
    d["b5"] = np.array([0.1])
    return mlp.pack(d)


def test_exact_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    in_dim, B = 7, 6
    p = _random_params(rng, in_dim)
    q = rng.uniform(-np.pi, np.pi, (B, in_dim))
    x = mlp.normalise(q, [-np.pi] * in_dim, [np.pi] * in_dim)
    t = rng.uniform(-1, 1, B)
    _, g = mlp.loss_and_grads(p, x, t, r=mlp.identity)
    # indices in every parameter block, including biases and the head
    idx = []
    for name, (o, s) in mlp.OFFSETS.items():
        n = int(np.prod(s))
        cols = s[1] if len(s) == 2 else n
        for _ in range(4):
            k = int(rng.integers(0, n))
            if name == "W1" and k % cols >= in_dim:  # padded input columns
                k -= k % cols
            idx.append(o + k)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = mlp.loss_and_grads(pp, x, t, r=mlp.identity)[0]
        lm = mlp.loss_and_grads(pm, x, t, r=mlp.identity)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_padded_inputs_have_zero_gradient():
    rng = np.random.default_rng(4)
    p = _random_params(rng, 5)
    q = rng.uniform(-1, 1, (8, 5))
    x = mlp.normalise(q, [-1] * 5, [1] * 5)
    _, g = mlp.loss_and_grads(p, x, rng.uniform(-1, 1, 8))
    W1 = mlp.unpack(g)["W1"]
    assert np.all(W1[:, 5:] == 0) and np.any(W1[:, :5] != 0)


def test_bf16_flavour_close_to_exact():
    rng = np.random.default_rng(5)
    p = _random_params(rng, 7)
    x = mlp.normalise(rng.uniform(-3, 3, (64, 7)), [-3.2] * 7, [3.2] * 7)
    y0, _ = mlp.forward(p, x, mlp.identity)
    y1, _ = mlp.forward(p, x, mlp.bf16)
    assert np.max(np.abs(y1 - y0)) < 0.05 * (1 + np.max(np.abs(y0)))
    assert np.max(np.abs(y1 - y0)) > 0  # the rounding points are real
