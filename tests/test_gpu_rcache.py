"""GPU radiance-cache MLP against the FP64 oracle (oracle/mlp.py): forward
outputs, the gradient of the relative L2 loss, one Adam step, and that
online training lowers the loss."""

import numpy as np
import pytest

import hdcodex_amd as hx
from oracle import mlp

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.12):
    return np.random.default_rng(seed).normal(scale=scale, size=mlp.PARAMS).astype(np.float32)


def _target(q):
    p, d = q["position"].astype(np.float64), q["direction"].astype(np.float64)
    return (1.0 + 0.5 * np.sin(2 * np.pi * p[:, 0]) + 0.3 * d[:, 1] + 0.2 * q["lane"]).astype(np.float32)


def test_inference_matches_oracle(tracer):
    params = _params(1)
    tracer.rc_set_params(params)
    q = mlp.random_queries(5000, 2)
    gpu = tracer.rc_infer(q)
    ref, _, _ = mlp.forward(params, q, bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(gpu - ref).max() / scale
    print("inference max error / max |y|", err)
    assert err < 2e-2
    assert np.abs(gpu - ref).mean() / np.abs(ref).mean() < 5e-3


def test_gradient_matches_fp64_oracle(tracer):
    params = _params(3)
    tracer.rc_set_params(params)
    q = mlp.random_queries(4096, 4, _target)
    loss, g = tracer.rc_gradient(q)
    ref_loss, ref = mlp.loss_and_gradient(params.astype(np.float64), q, bf16=True)
    assert abs(loss - ref_loss) / ref_loss < 1e-2
    for i, (a, b) in enumerate(zip(mlp.OFFSETS[:-1], mlp.OFFSETS[1:])):
        gg, rr = g[a:b].astype(np.float64), ref[a:b]
        cos = float(gg @ rr / (np.linalg.norm(gg) * np.linalg.norm(rr)))
        rel = float(np.linalg.norm(gg - rr) / np.linalg.norm(rr))
        print(f"layer {i}: cosine {cos:.5f} relative error {rel:.4f}")
        assert cos > 0.999 and rel < 4e-2


def test_adam_step_matches_oracle(tracer):
    params = _params(5)
    tracer.rc_set_params(params)
    q = mlp.random_queries(2048, 6, _target)
    _, g = tracer.rc_gradient(q)
    tracer.rc_set_params(params)
    tracer.rc_train(q, steps=1)
    new = tracer.rc_params()
    ref, _, _ = mlp.adam(params.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1)
    assert np.allclose(new, ref, rtol=1e-5, atol=1e-6)


def test_training_lowers_the_loss(tracer):
    tracer.rc_set_params(_params(7, 0.08))
    q = mlp.random_queries(65536, 8, _target)
    first = tracer.rc_train(q, steps=4)
    for _ in range(40):
        last = tracer.rc_train(q, steps=4)
    print("loss", first, "->", last)
    assert last < 0.2 * first
    test = mlp.random_queries(4096, 9, _target)
    y = tracer.rc_infer(test)[np.arange(len(test)), test["lane"].astype(int)]
    rel = np.mean((y - test["target"]) ** 2 / (test["target"] ** 2 + 0.01))
    print("held-out relative L2", rel)
    assert rel < 0.05
