"""Pins the FP64 MLP oracle: its gradient against central finite
differences of its own loss, Adam against the published update rule, and
the encoding against closed forms."""

import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = rng.normal(scale=0.15, size=mlp.PARAMS)
    q = mlp.random_queries(48, 1, lambda q: 0.5 + q["position"][:, 0])
    loss, grad = mlp.loss_and_gradient(params, q, bf16=False)
    # Relative L2 with the denominator held constant: differentiate the loss
    # with that denominator frozen at the current outputs.
    y0, _, _ = mlp.forward(params, q, bf16=False)
    den = y0[np.arange(len(q)), q["lane"].astype(int)] ** 2 + 0.01

    def frozen_loss(p):
        y, _, _ = mlp.forward(p, q, bf16=False)
        yk = y[np.arange(len(q)), q["lane"].astype(int)]
        return float(np.mean((yk - q["target"]) ** 2 / den))

    assert abs(frozen_loss(params) - loss) < 1e-12
    for i in rng.choice(mlp.PARAMS, size=40, replace=False):
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (frozen_loss(p1) - frozen_loss(p2)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


def test_adam_first_step_is_sign_times_lr():
    # With zero moments, bias correction makes the first Adam step
    # lr * g / (|g| + eps) (Kingma & Ba 2015).
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p, m, v = mlp.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, lr=1e-2, eps=1e-8)
    assert np.allclose(p, -1e-2 * g / (np.abs(g) + 1e-8))


def test_encoding_closed_forms():
    q = np.zeros(1, dtype=mlp.QUERY_DTYPE)
    q["position"] = [0.25, 0.0, 0.5]
    q["direction"] = [1.0, 0.0, -1.0]
    q["lambda0"] = 580.0
    q["roughness"] = 0.3
    q["albedo"] = 0.7
    f = mlp.encode(q)[0]
    assert np.isclose(f[0], np.sin(np.pi * 0.25)) and np.isclose(f[1], np.cos(np.pi * 0.25))
    assert np.isclose(f[2], np.sin(2 * np.pi * 0.25)) and np.isclose(f[3], np.cos(2 * np.pi * 0.25))
    assert np.isclose(f[24], 0.0) and np.isclose(f[25], 1.0)
    assert np.isclose(f[72], np.sin(np.pi * 1.0)) and np.isclose(f[73], np.cos(np.pi * 1.0))
    assert np.isclose(f[96], np.sin(0.0))
    assert np.allclose(f[108:112], [0.5, 0.3, 0.7, 1.0]) and np.all(f[112:] == 0)


def test_round_bf16_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 65504.0])
    r = mlp.round_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5 and r[4] == 65536.0
