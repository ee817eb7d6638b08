"""GPU: the device BVH builder and refit.

* The device build equals the serial CPU run of the same steps node for node
  (after renumbering both trees in breadth-first order: the device hands out
  node ids with an atomic counter) and slot for slot.
* Renders do not depend on the builder: closest hits are the smallest
  (t, primitive) pair and any-hit candidates are visited in that order, so
  images from the three builders are bit-identical.
* Refit: hdcx_set_scene with moved vertices and the same triangles refits
  the BVH; hits on the moved scene equal the oracle's bit for bit.
"""

import ctypes

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import scene_rays

pytestmark = pytest.mark.gpu

LEAF = 0x80000000
EMPTY = 0xFFFFFFFF


def _canonical(nodes):
    """Nodes in breadth-first order from the root with inner child ids
    renumbered to that order."""
    order = [0]
    index = {0: 0}
    i = 0
    children = nodes["child"]
    while i < len(order):
        for c in children[order[i]]:
            c = int(c)
            if c != EMPTY and not c & LEAF:
                index[c] = len(order)
                order.append(c)
        i += 1
    assert len(order) == len(nodes), "unreachable nodes"
    out = nodes[np.array(order)].copy()
    ch = out["child"].astype(np.int64)
    inner = (ch != EMPTY) & ((ch & LEAF) == 0)
    ch[inner] = [index[int(c)] for c in ch[inner]]
    out["child"] = ch.astype(np.uint32)
    return out


@pytest.mark.parametrize("clusters", [0, 2048])
@pytest.mark.parametrize("name,args", [("cornell", (0,)), ("materials", (0,)), ("heightfield", (1, 256, 64)),
                                       ("heightfield", (0, 1024, 64))])
def test_device_lbvh_equals_serial_run(name, args, clusters):
    scene = hx.ProceduralScene(name, *args)
    dinfo, dnodes, dslots = hx.build_bvh(scene, hx.BVH_DEVICE_LBVH, clusters)
    sinfo, snodes, sslots = hx.build_bvh(scene, hx.BVH_SERIAL_LBVH, clusters)
    print(f"{name} {scene.triangle_count} tris: device {dinfo['build_ms']:.1f} ms, serial {sinfo['build_ms']:.1f} ms, "
          f"{dinfo['node_count']} nodes, stack {dinfo['max_stack']}, top clusters {dinfo['top_clusters']}")
    for k in ("node_count", "slot_count", "max_stack", "depth", "top_clusters"):
        assert dinfo[k] == sinfo[k], k
    assert np.array_equal(dslots.view(np.uint32), sslots.view(np.uint32))
    a, b = _canonical(dnodes), _canonical(snodes)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name,args", [("triangle", (0,)), ("materials", (0,)), ("heightfield", (1, 128, 64))])
def test_renders_do_not_depend_on_the_builder(name, args):
    scene = hx.ProceduralScene(name, *args)
    cam = scene.camera(64, 32)
    images = {}
    for builder in (hx.BVH_HOST_SAH, hx.BVH_DEVICE_LBVH, hx.BVH_SERIAL_LBVH):
        t = hx.PathTracer(0)
        try:
            t.set_bvh_options(builder, 2048, True)
            t.set_scene(scene)
            if scene.triangle_count >= 2:
                assert t.bvh_info()["builder"] == builder
            images[builder] = t.render(cam, 64, 32, 0, 8, 4)
        finally:
            t.close()
    ref = images[hx.BVH_HOST_SAH]
    for builder, img in images.items():
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), builder


def _moved(scene):
    """A copy of the scene view with every vertex displaced (same triangles)."""
    v = hx.SceneView.from_buffer_copy(scene.view)
    pos = scene.arrays()["positions"]
    pos[:, 1] = pos[:, 1] * 1.25 + 0.8 * np.sin(pos[:, 0] * 0.2)
    pos[:, 0] += 0.3 * np.cos(pos[:, 2] * 0.15)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    v.positions = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return v, pos


@pytest.mark.parametrize("builder", [hx.BVH_HOST_SAH, hx.BVH_DEVICE_LBVH])
def test_refit_hits_match_oracle(builder):
    scene = hx.ProceduralScene("heightfield", 2, 128, 64)
    t = hx.PathTracer(0)
    try:
        t.set_bvh_options(builder, 2048, True)
        t.set_scene(scene)
        assert t.bvh_info()["refits"] == 0
        view, keep = _moved(scene)
        t.set_scene(view)
        info = t.bvh_info()
        assert info["refits"] == 1 and info["builder"] == builder
        cam = scene.camera(160, 90)
        lo, hi = keep[:, :3].min(axis=0), keep[:, :3].max(axis=0)
        rays = scene_rays(cam, 20000, 3, inside=(lo, hi))
        gpu = t.trace_closest(rays)
        ref, _ = oracle.OracleScene(view).trace_closest(rays)
        hit = ref["primitive"] != EMPTY
        print(f"refit ({info['build_ms']:.2f} ms): {int(hit.sum())} hits")
        assert hit.sum() > 5000
        assert np.array_equal(gpu["primitive"], ref["primitive"])
        for k in ("t", "u", "v"):
            assert np.array_equal(gpu[k][hit].view(np.uint32), ref[k][hit].view(np.uint32))
        # A different index buffer rebuilds.
        t.set_scene(hx.ProceduralScene("cornell", 0))
        assert t.bvh_info()["refits"] == 0
    finally:
        t.close()
