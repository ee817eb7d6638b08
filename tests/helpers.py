"""Shared helpers of the test-suite (numerics of the parity checks)."""

from __future__ import annotations

import numpy as np


def rel_mse(x: np.ndarray, y: np.ndarray) -> float:
    """Mean over pixels and RGB of (x - y)^2 / (y^2 + 0.01); y is the oracle."""
    x = np.asarray(x, dtype=np.float64)[..., :3]
    y = np.asarray(y, dtype=np.float64)[..., :3]
    return float(np.mean((x - y) ** 2 / (y * y + 0.01)))


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in float32 ulps between finite float32 arrays."""
    ai = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def scene_rays(camera, n: int, seed: int, spread: float = 1.0, inside: np.ndarray | None = None):
    """Seeded query rays: half from the camera through jittered image points,
    half from random points (inside a box when given) in random directions."""
    rng = np.random.default_rng(seed)
    cam = np.array([list(camera.origin), list(camera.lower_left), list(camera.horizontal), list(camera.vertical)],
                   dtype=np.float64)
    k = n // 2
    uv = rng.random((k, 2))
    target = cam[1] + uv[:, :1] * cam[2] + uv[:, 1:] * cam[3]
    d1 = target - cam[0]
    o1 = np.repeat(cam[:1], k, axis=0)
    if inside is None:
        lo, hi = cam[0] - spread, cam[0] + spread
    else:
        lo, hi = inside
    o2 = lo + (hi - lo) * rng.random((n - k, 3))
    d2 = rng.normal(size=(n - k, 3))
    o = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = np.empty((n, 8), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = 1e4
    return r
