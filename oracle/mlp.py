"""ORACLE — FP64 restatement of the radiance-cache MLP (test infrastructure).

Same encoding, layer layout and loss as the HIP kernels
(hdcodex_amd/csrc/rcache.hip), computed in float64. ``round_bf16`` lets a
check apply the kernels' BF16 rounding points (inputs, weights, hidden
activations) so that what remains is accumulation-order error.

Layout: W0 [64][128], W1..W4 [64][64], W5 [3][64] row-major, concatenated
(24768 parameters); y = W5 relu(W4 ... relu(W0 x)).
Loss per sample (relative L2 on the sample's lane k, denominator held
constant): (y_k - t)^2 / (y_k^2 + 0.01), averaged over samples.
"""

from __future__ import annotations

import numpy as np

SHAPES = [(64, 128), (64, 64), (64, 64), (64, 64), (64, 64), (3, 64)]
OFFSETS = np.cumsum([0] + [a * b for a, b in SHAPES])
PARAMS = int(OFFSETS[-1])
QUERY_DTYPE = np.dtype([("position", np.float32, 3), ("direction", np.float32, 3), ("lambda0", np.float32),
                        ("roughness", np.float32), ("albedo", np.float32), ("lane", np.uint32),
                        ("path", np.uint32), ("target", np.float32)])


def round_bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (b.astype(np.uint32)).view(np.float32).astype(np.float64)


def split(params: np.ndarray) -> list[np.ndarray]:
    p = np.asarray(params, dtype=np.float64)
    return [p[OFFSETS[i]:OFFSETS[i + 1]].reshape(SHAPES[i]) for i in range(len(SHAPES))]


def encode(q: np.ndarray) -> np.ndarray:
    """Frequency encoding (N, 128): sin/cos(2^k pi p) for 12 bands of each
    position axis, sin/cos(2^k pi (d+1)/2) for 6 bands of each direction
    axis, then (lambda0 - 380)/400, roughness, albedo, 1, zeros."""
    n = len(q)
    f = np.zeros((n, 128))
    p = q["position"].astype(np.float64)
    d = (q["direction"].astype(np.float64) + 1.0) * 0.5
    for a in range(3):
        for k in range(12):
            arg = np.pi * (2.0 ** k) * p[:, a]
            f[:, a * 24 + 2 * k] = np.sin(arg)
            f[:, a * 24 + 2 * k + 1] = np.cos(arg)
        for k in range(6):
            arg = np.pi * (2.0 ** k) * d[:, a]
            f[:, 72 + a * 12 + 2 * k] = np.sin(arg)
            f[:, 72 + a * 12 + 2 * k + 1] = np.cos(arg)
    f[:, 108] = (q["lambda0"].astype(np.float64) - 380.0) / 400.0
    f[:, 109] = q["roughness"]
    f[:, 110] = q["albedo"]
    f[:, 111] = 1.0
    return f


def forward(params: np.ndarray, q: np.ndarray, bf16: bool = True):
    """Returns outputs (N, 3) and the list of layer inputs/activations."""
    ws = split(params)
    if bf16:
        ws = [round_bf16(w) for w in ws]
    x = encode(q)
    if bf16:
        x = round_bf16(x)
    acts = [x]
    h = x
    for i in range(5):
        h = np.maximum(h @ ws[i].T, 0.0)
        if bf16:
            h = round_bf16(h)
        acts.append(h)
    y = h @ ws[5].T
    return y, acts, ws


def loss_and_gradient(params: np.ndarray, q: np.ndarray, bf16: bool = True):
    """Mean relative L2 loss and its gradient (flat, float64)."""
    y, acts, ws = forward(params, q, bf16)
    n = len(q)
    lane = q["lane"].astype(np.int64)
    yk = y[np.arange(n), lane]
    den = yk * yk + 0.01
    diff = yk - q["target"].astype(np.float64)
    loss = float(np.mean(diff * diff / den))
    dy = np.zeros_like(y)
    dy[np.arange(n), lane] = 2.0 * diff / den / n
    grads = [None] * 6
    grads[5] = dy.T @ acts[5]
    delta = (dy @ ws[5]) * (acts[5] > 0)
    for i in range(4, -1, -1):
        grads[i] = delta.T @ acts[i]
        if i > 0:
            delta = (delta @ ws[i]) * (acts[i] > 0)
    return loss, np.concatenate([g.reshape(-1) for g in grads])


def adam(params, m, v, grad, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8):
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def random_queries(n: int, seed: int, target_fn=None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q = np.zeros(n, dtype=QUERY_DTYPE)
    q["position"] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    q["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q["lambda0"] = rng.uniform(380, 780, n)
    q["roughness"] = rng.random(n)
    q["albedo"] = rng.random(n)
    q["lane"] = rng.integers(0, 3, n)
    q["path"] = np.arange(n)
    if target_fn is not None:
        q["target"] = target_fn(q)
    return q
