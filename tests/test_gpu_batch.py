"""GPU: the batched multi-scene renderer.

Every lane path of a batch runs the same kernels on the same seeds as in a
single-scene render, and each pixel's samples are summed in the same order,
so a scene's film from the batch equals the single-tracer film bit for bit,
whatever the wave / tile / grid sizes and whether the per-wave graph is
replayed. Against the FP64 oracle every scene stays within rel-MSE 1e-4.
"""

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import rel_mse

pytestmark = pytest.mark.gpu

# (name, args, width, height, spp)
SCENES = [
    ("cornell", (0,), 48, 40, 8),
    ("materials", (0,), 64, 32, 4),
    ("mixed", (3, 3000, 64), 40, 40, 6),
    ("heightfield", (1, 64, 64), 56, 32, 3),
    ("mixed_lights", (5, 1500, 32), 32, 24, 5),
]


@pytest.fixture(scope="module")
def scenes():
    return [(hx.ProceduralScene(n, *a), w, h, spp) for n, a, w, h, spp in SCENES]


def _single(scene, w, h, spp, sample_index=0):
    t = hx.PathTracer(0)
    try:
        t.set_bvh_options(hx.BVH_DEVICE_LBVH, 2048, True)
        t.set_scene(scene)
        cam = scene.camera(w, h)
        if sample_index:
            t.render(cam, w, h, 0, 8, sample_index)
        return t.render(cam, w, h, sample_index, 8, spp)
    finally:
        t.close()


def _batch(scenes, tuning=None, sample_index=0):
    b = hx.BatchRenderer(0)
    try:
        for i, (scene, w, h, spp) in enumerate(scenes):
            assert b.add_scene(scene) == i
            b.set_view(i, scene.camera(w, h), w, h, spp)
        if tuning:
            b.set_tuning(*tuning)
        if sample_index:
            # Samples [0, sample_index) first: each scene renders sample_index
            # samples per call only when its view asks for that many.
            for i, (scene, w, h, spp) in enumerate(scenes):
                b.set_view(i, scene.camera(w, h), w, h, sample_index)
            b.render(0, 8)
            for i, (scene, w, h, spp) in enumerate(scenes):
                b.set_view(i, scene.camera(w, h), w, h, spp)
        b.render(sample_index, 8)
        films = [b.read_film(i) for i in range(len(scenes))]
        return films, b.stats()
    finally:
        b.close()


@pytest.mark.parametrize("tuning", [None, (3 << 20, 4096, 2048, True), (3 * 8 * 300, 97, 64, True),
                                    (3 * 8 * 300, 97, 64, False)])
def test_batch_equals_single_scene_renders(scenes, tuning):
    films, stats = _batch(scenes, tuning)
    print("batch stats", stats)
    assert stats["pixel_samples"] == sum(w * h * spp for _, w, h, spp in scenes)
    if tuning and tuning[0] < 10000:
        assert stats["waves"] > 10
    if tuning is None or tuning[3]:
        assert stats["graph_launches"] == stats["waves"]
    for (scene, w, h, spp), film in zip(scenes, films):
        ref = _single(scene, w, h, spp)
        assert film.shape == ref.shape
        assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("tuning", [None, (3 * 8 * 300, 97, 64, True)])
def test_opaque_batch_equals_single_scene_renders(scenes, tuning):
    """A batch whose scenes are all opaque (Cornell box, the materials
    scene, the heightfield) runs extend as the lane-refill kernel over the
    scene table, each lane holding its own scene: films equal the
    single-scene renders bit for bit."""
    opaque = [scenes[0], scenes[1], scenes[3]]
    assert all(scene.arrays()["opaque"] for scene, _, _, _ in opaque)
    films, _ = _batch(opaque, tuning)
    for (scene, w, h, spp), film in zip(opaque, films):
        ref = _single(scene, w, h, spp)
        assert np.array_equal(film.view(np.uint32), ref.view(np.uint32))


def test_batch_progressive_mean(scenes):
    films, _ = _batch(scenes[:3], sample_index=5)
    for (scene, w, h, spp), film in zip(scenes[:3], films):
        ref = _single(scene, w, h, spp, sample_index=5)
        assert np.max(np.abs(film - ref)) < 1e-5


def test_batch_matches_fp64_oracle(scenes):
    films, _ = _batch(scenes)
    for (scene, w, h, spp), film in zip(scenes, films):
        ref = oracle.OracleScene(scene.view).render(scene.camera(w, h), w, h, 0, 8, spp, precision=64) / spp
        err = rel_mse(film, ref)
        print(f"scene {scene.triangle_count} tris: rel-MSE vs FP64 oracle {err:.3e}")
        assert err < 1e-4


def test_autotune_picks_and_caches_a_tuning(scenes, tmp_path):
    cache = tmp_path / "tuning.txt"
    b = hx.BatchRenderer(0)
    try:
        for i, (scene, w, h, spp) in enumerate(scenes):
            b.add_scene(scene)
            b.set_view(i, scene.camera(w, h), w, h, spp)
        first = b.autotune(8, str(cache))
        assert first["wave_paths"] >= 3 and first["tile_pixels"] > 0 and first["max_blocks"] > 0
        assert cache.read_text().count("\n") == 1
        b.set_tuning(999, 7, 3, False)
        again = b.autotune(8, str(cache))
        assert again == first
        assert cache.read_text().count("\n") == 1
        b.render(0, 8)
        assert b.stats()["pixel_samples"] == sum(w * h * spp for _, w, h, spp in scenes)
    finally:
        b.close()


def test_batch_with_small_waves_is_reproducible(scenes):
    """Small waves replayed from one graph, whose queues are filled wave by
    wave in atomic order (a different order on every run): every
    repetition equals the single-scene renders bit for bit."""
    refs = [_single(scene, w, h, spp) for scene, w, h, spp in scenes]
    for rep in range(8):
        films, _ = _batch(scenes, (3 * 8 * 300, 97, 64, True))
        for film, ref in zip(films, refs):
            assert np.array_equal(film.view(np.uint32), ref.view(np.uint32)), rep
