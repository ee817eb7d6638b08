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


OPACITY_OFFSET = 36  # transmissionOpacityIor.y in the 368-byte material record


def cutout_view(scene):
    """The scene's view with every other material at opacity 0.5 and the
    opaque flag cleared (a scene with cut-outs). Returns (view, buffer); keep
    the buffer alive while the view is used."""
    import hdcodex_amd as hx

    v = hx.SceneView.from_buffer_copy(scene.view)
    mats = np.array(scene.arrays()["materials"], dtype=np.uint8, copy=True)
    op = mats[:, OPACITY_OFFSET:OPACITY_OFFSET + 4].copy().view(np.float32).reshape(-1)
    op[::2] = 0.5
    mats[:, OPACITY_OFFSET:OPACITY_OFFSET + 4] = op.view(np.uint8).reshape(-1, 4)
    mats = np.ascontiguousarray(mats)
    v.materials = mats.ctypes.data
    v.opaque = 0
    return v, mats


def cpu_threads() -> int:
    """Worker threads for the CPU oracle: the scheduler affinity, capped by
    OMP_NUM_THREADS when set (the host share a job is granted)."""
    import os

    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def texture_mean_luminance(view, i: int) -> float:
    """Mean linear luminance of texture i of a view (numpy, FP64)."""
    import ctypes

    t = view.textures[i]
    n = t.width * t.height * 4
    if t.format == 2:
        x = np.ctypeslib.as_array(ctypes.cast(t.texels, ctypes.POINTER(ctypes.c_float)), shape=(n,)).astype(np.float64)
    else:
        b = np.ctypeslib.as_array(ctypes.cast(t.texels, ctypes.POINTER(ctypes.c_uint8)), shape=(n,)) / 255.0
        x = np.where(b <= 0.04045, b / 12.92, ((b + 0.055) / 1.055) ** 2.4) if t.format == 1 else b
    rgb = np.clip(x.reshape(-1, 4)[:, :3], 0, None)
    return float((rgb @ np.array([0.2126, 0.7152, 0.0722])).mean())


def with_emission_texture(scene, slot: int):
    """A copy of the scene's view whose emissive materials take their
    emission from texture slot `slot` (a plain index or a UDIM set), with no
    per-material powers: the device and the oracle derive them from the
    records. Returns (view, keep-alive buffers, expected powers)."""
    import ctypes

    import hdcodex_amd as hx

    v = scene.view
    view = hx.SceneView()
    ctypes.memmove(ctypes.byref(view), ctypes.byref(v), ctypes.sizeof(view))
    mats = np.frombuffer(ctypes.string_at(v.materials, 368 * v.material_count), dtype=np.uint8).copy()
    f = mats.view(np.float32).reshape(-1, 92)
    u = mats.view(np.uint32).reshape(-1, 92)
    emitters = f[:, 11] > 0
    u[emitters, 71] = slot
    if slot & 0x80000000:
        tiles = [(slot & 0xFF) + k for k in range(bin((slot >> 8) & 0x7FFFFF).count("1"))]
    else:
        tiles = [slot]
    lum = np.float32(np.mean([np.float32(texture_mean_luminance(v, t)) for t in tiles]))
    power = np.where(emitters, np.maximum(f[:, 11], 0) * lum, 0.0).astype(np.float32)
    view.materials = mats.ctypes.data
    view.material_emission_power = None
    return view, mats, power
