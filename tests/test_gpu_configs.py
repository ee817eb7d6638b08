"""GPU: the benchmark workloads at full size against the FP64 oracle.

* Heightfield frame: 1024 x 1024 vertex grid (2,093,058 triangles), 2048 x
  1024 HDR dome, 1920 x 1080, 16 spp x 3 lanes, 8 bounces, with the host SAH
  and the device LBVH trees; rel-MSE < 1e-4 against the FP64 oracle on every
  8th row and column (32,400 pixels) at the same seeds.
* 64-scene batch: the seeded mixed-material scenes of the benchmark (1e3 to
  5e5 triangles, 128^2 to 1024^2 pixels, 8 to 64 spp) in one batch replayed
  from the captured per-wave graph; every scene's 32 pixels (4 rows x 8
  columns) < 1e-4 against the FP64 oracle.
* Radiance cache on the heightfield: after online training, the cached
  frame (paths end at their second vertex) has a lower rel-MSE than the
  plain frame of the same sample count against the committed FP64
  references (4096 spp on every 8th row and column, 32,768 spp on every
  32nd), at least 2x lower against the first.
"""

from pathlib import Path

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import cpu_threads, rel_mse

pytestmark = pytest.mark.gpu

W, H, SPP, BOUNCES = 1920, 1080, 16, 8
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def heightfield():
    scene = hx.ProceduralScene("heightfield", 0, 1024, 2048)
    return scene, scene.camera(W, H)


@pytest.mark.parametrize("builder", [hx.BVH_HOST_SAH, hx.BVH_DEVICE_LBVH], ids=["host_sah", "device_lbvh"])
def test_heightfield_frame_matches_fp64_oracle(heightfield, builder):
    scene, cam = heightfield
    t = hx.PathTracer(0)
    try:
        t.set_bvh_options(builder, 2048, True)
        t.set_scene(scene)
        img = t.render(cam, W, H, 0, BOUNCES, SPP)
    finally:
        t.close()
    assert img.shape == (H, W, 4) and np.all(np.isfinite(img)) and np.all(img[..., 3] == 1.0)
    rows = np.arange(0, H, 8, dtype=np.uint32)
    ref = oracle.OracleScene(scene.view).render(cam, W, H, 0, BOUNCES, SPP, precision=64, rows=rows, col_step=8,
                                                threads=cpu_threads()) / SPP
    err = rel_mse(img[rows.astype(np.int64)][:, ::8], ref)
    print(f"heightfield 1080p x 16 spp, builder {builder}: rel-MSE vs FP64 oracle {err:.3e} on {ref.shape[0] * ref.shape[1]} px")
    assert err < 1e-4


def batch_specs():
    """The benchmark's 64-scene batch (bench.py batch_workload): scene index,
    triangle target, square resolution, spp."""
    rng = np.random.default_rng(2024)
    specs = []
    for i in range(64):
        tris = int(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        res = int(rng.choice([128, 256, 384, 512, 640, 768, 896, 1024]))
        spp = int(rng.integers(8, 65))
        specs.append((i, tris, res, spp))
    return specs


def test_batch_64_scenes_graph_replay_matches_fp64_oracle():
    specs = batch_specs()
    b = hx.BatchRenderer(0)
    scenes = []
    try:
        for i, tris, res, spp in specs:
            scene = hx.ProceduralScene("mixed", i, tris, 256)
            scenes.append(scene)
            b.add_scene(scene, hx.BVH_DEVICE_LBVH, 2048)
            b.set_view(i, scene.camera(res, res), res, res, spp)
        # The autotuner's choice on this workload: 25.2 M lane paths per wave,
        # 16,384-pixel tiles, 4,096 blocks per launch; replayed graph.
        b.set_tuning(3 << 23, 16384, 4096, True)
        b.render(0, BOUNCES)
        stats = b.stats()
        films = [b.read_film(i) for i in range(len(specs))]
    finally:
        b.close()
    assert stats["graph_launches"] == stats["waves"] >= 2
    errs = []
    threads = cpu_threads()
    for (i, tris, res, spp), scene in zip(specs, scenes):
        rows = np.linspace(res // 8, res - res // 8, 4).astype(np.uint32)
        step = res // 8
        cam = scene.camera(res, res)
        ref = oracle.OracleScene(scene.view).render(cam, res, res, 0, BOUNCES, spp, precision=64, rows=rows,
                                                    col_step=step, threads=threads) / spp
        x = films[i][rows.astype(np.int64)][:, ::step]
        assert np.all(np.isfinite(x))
        errs.append(rel_mse(x, ref))
    print(f"64 scenes, {stats['waves']} waves: rel-MSE vs FP64 max {max(errs):.3e} median {np.median(errs):.3e}")
    assert max(errs) < 1e-4


def _score(frame: np.ndarray, ref) -> float:
    rows, cols = ref["rows"].astype(np.int64), ref["cols"].astype(np.int64)
    y = 0.5 * (ref["half_a"] + ref["half_b"])
    return rel_mse(frame[rows][:, cols], y)


def test_trained_cache_beats_plain_at_equal_samples(heightfield):
    """What the trained network gains over the method without it: the same
    16 spp frame with paths ending into the cache at their second vertex
    scores better against both FP64 references than the plain frame."""
    scene, cam = heightfield
    refs = {n: np.load(GOLDEN / f"c1_reference_{n}spp.npz") for n in (4096, 32768)}
    t = hx.PathTracer(0)
    try:
        t.set_bvh_options(hx.BVH_HOST_SAH, 2048, True)
        t.set_scene(scene)
        plain = t.render(cam, W, H, 0, BOUNCES, SPP)
        t.set_radiance_cache(True, train_samples=65536, terminate_bounce=1, train_steps=4, learning_rate=1e-2, seed=1,
                             ema_decay=0.99)
        for i in range(50):
            t.render(cam, W, H, 1000 + 2 * i, BOUNCES, 2, readback=False)
        cached = t.render(cam, W, H, 0, BOUNCES, SPP)
        stats = t.rc_stats()
    finally:
        t.close()
    assert stats["queries"] > 0 and np.isfinite(stats["last_loss"])
    for n, ref in refs.items():
        p, c = _score(plain, ref), _score(cached, ref)
        print(f"{n}-spp reference: plain {p:.4f}, cache {c:.4f}, ratio {p / c:.2f}")
        assert c < p
    # On the 32,400 pixels of the 4096-spp reference the plain frame's error
    # is dominated by rare bright paths after the second vertex, which the
    # cache ends: measured 1.05 against 0.24.
    assert _score(cached, refs[4096]) * 2 < _score(plain, refs[4096])
