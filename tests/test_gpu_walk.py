"""GPU: subsurface random walks deferred to their own kernel.

Shade stops a path at a subsurface walk and queues it; the walk kernel runs
the walks; shade resumes the queued paths from their unchanged state. The
images must equal, bit for bit, those of the single-kernel form in which
shade runs each walk itself (instrumentation bit 2), with and without the
radiance cache, for one-level and instanced scenes.
"""

import numpy as np
import pytest

import hdcodex_amd as hx

pytestmark = pytest.mark.gpu

TIMED, INLINE_WALKS = 1, 4
SCENES = [("materials", (0,)), ("mixed", (3, 20000, 64)), ("instanced_small", (1, 30, 64))]


def _render(scene, w, h, spp, flags, cache=False):
    t = hx.PathTracer(0)
    try:
        t.set_scene(scene)
        if cache:
            t.set_radiance_cache(True, train_samples=4096, terminate_bounce=1, seed=3, spread_threshold=0.05)
        t.set_instrumentation(TIMED | flags)
        img = t.render(scene.camera(w, h), w, h, 0, 8, spp)
        return img, t.kernel_stats()
    finally:
        t.close()


@pytest.mark.parametrize("kind,args", SCENES, ids=[s[0] for s in SCENES])
def test_deferred_walks_equal_inline_walks(kind, args):
    scene = hx.ProceduralScene(kind, *args)
    deferred, ds = _render(scene, 96, 64, 4, 0)
    inline, ins = _render(scene, 96, 64, 4, INLINE_WALKS)
    print(kind, "walk launches", ds["walk"]["launches"], "walk ms", ds["walk"]["ms"])
    assert ds["walk"]["launches"] > 0 and ins["walk"]["launches"] == 0
    assert np.all(np.isfinite(deferred))
    assert np.array_equal(deferred.view(np.uint32), inline.view(np.uint32))


def test_deferred_walks_with_radiance_cache():
    scene = hx.ProceduralScene("mixed", 5, 20000, 64)
    deferred, _ = _render(scene, 64, 48, 2, 0, cache=True)
    inline, _ = _render(scene, 64, 48, 2, INLINE_WALKS, cache=True)
    assert np.array_equal(deferred.view(np.uint32), inline.view(np.uint32))


def test_scene_without_subsurface_runs_no_walk_kernel(tracer):
    scene = hx.ProceduralScene("cornell", 0)
    tracer.set_scene(scene)
    tracer.reset_kernel_stats()
    tracer.set_instrumentation(TIMED)
    try:
        tracer.render(scene.camera(32, 32), 32, 32, 0, 8, 2)
        assert tracer.kernel_stats()["walk"]["launches"] == 0
    finally:
        tracer.set_instrumentation(0)
