"""GPU: one radiance cache per scene in the batch renderer.

Queries of all scenes share one queue; they are bucketed by scene and each
bucket is answered by its scene's network. A scene's film therefore does not
depend on which other scenes share the batch: with the same initial weights
(seed + scene index) and the same training choice (per scene), the first
render of a scene alone equals its render next to other scenes, bit for
bit, with and without graph replay. Training runs per scene after the
render; later frames stay finite and close to the plain films.
"""

import numpy as np
import pytest

import hdcodex_amd as hx

pytestmark = pytest.mark.gpu

SCENES = [("mixed", (3, 3000, 64), 48, 40, 4), ("materials", (0,), 40, 24, 3), ("cornell", (0,), 32, 32, 4)]
CACHE = dict(train_samples=8192, terminate_bounce=1, train_steps=2, ema_decay=0.9)


@pytest.fixture(scope="module")
def scenes():
    return [(hx.ProceduralScene(n, *a), w, h, spp) for n, a, w, h, spp in SCENES]


def _batch(entries, seed, graph=True, frames=1, cache=True):
    b = hx.BatchRenderer(0)
    try:
        for i, (s, w, h, spp) in enumerate(entries):
            b.add_scene(s)
            b.set_view(i, s.camera(w, h), w, h, spp)
        b.set_tuning(1 << 16, 256, 256, graph)
        if cache:
            b.set_radiance_cache(True, seed=seed, **CACHE)
        films = []
        for f in range(frames):
            b.render(f * 64, 8)
            films.append([b.read_film(i) for i in range(len(entries))])
        return films, (b.rc_stats() if cache else None)
    finally:
        b.close()


def test_scene_film_independent_of_batch_neighbours(scenes):
    together, stats = _batch(scenes, seed=10)
    print("rc stats", stats)
    assert stats["queries"] > 0 and stats["scenes_trained"] == len(scenes)
    for i, entry in enumerate(scenes):
        alone, _ = _batch([entry], seed=10 + i)
        assert np.all(np.isfinite(alone[0][0]))
        assert np.array_equal(together[0][i].view(np.uint32), alone[0][0].view(np.uint32)), f"scene {i}"


def test_cache_films_equal_with_and_without_graph(scenes):
    a, _ = _batch(scenes, seed=4, graph=True)
    b, _ = _batch(scenes, seed=4, graph=False)
    for i in range(len(scenes)):
        assert np.array_equal(a[0][i].view(np.uint32), b[0][i].view(np.uint32))


def test_trained_caches_stay_close_to_plain(scenes):
    cached, stats = _batch(scenes, seed=1, frames=6)
    plain, _ = _batch(scenes, seed=1, frames=1, cache=False)
    print("rc stats", stats)
    assert stats["adam_steps"] == 6 * len(scenes) * CACHE["train_steps"]
    for i in range(len(scenes)):
        c, p = cached[-1][i][..., :3], plain[0][i][..., :3]
        assert np.all(np.isfinite(c))  # (sRGB of spectral samples may be slightly negative, cached or not)
        ratio = float(c.mean() / p.mean())
        print(SCENES[i][0], "mean ratio cache / plain", ratio, "min", float(c.min()), float(p.min()))
        assert 0.5 < ratio < 1.5
