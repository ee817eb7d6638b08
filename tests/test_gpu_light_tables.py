"""GPU: the light tables hdcx_set_scene builds on the device against the
oracle's FP64 restatement of the reference's host builds.

* Emissive-triangle table (vulkan_path_tracer.cpp:2944-2993) and lat-long
  dome importance tables (:1998-2066): the device sums run in a fixed tree
  order, the reference's (and the oracle's) sequentially, so the FP64
  values differ by a few FP64 ulps before the FP32 store. Bar: CDF entries
  within 2^-23 absolute (one FP32 ulp at 1.0), pdf entries within 4 FP32
  ulps relative (8 x 2^-24), areas bit-exact; zero entries stay zero.
* Emission powers derived from the records when the view carries none
  (colour, plain emission texture, UDIM set) match the oracle's rule.
* A dome with no positive luminance falls back to uniform sampling.
* A view that brings host-built tables (the reference's packing: emissive
  table given, importance RGBA32F texture in the dome's importance slot)
  is sampled from those tables unchanged.
* Images with the device-built tables: rel-MSE < 1e-4 against the FP64
  oracle at the same seeds.
"""

import ctypes

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import cpu_threads, rel_mse, with_emission_texture

pytestmark = pytest.mark.gpu

MISSING = 0xFFFFFFFF
CDF_ATOL = 2.0 ** -23
PDF_RTOL = 8 * 2.0 ** -24


def copy_view(view) -> hx.SceneView:
    out = hx.SceneView()
    ctypes.memmove(ctypes.byref(out), ctypes.byref(view), ctypes.sizeof(out))
    return out


def check_emissive(dev: np.ndarray, ref: np.ndarray) -> None:
    assert dev.shape == ref.shape
    assert np.all(np.abs(dev[:, 0].astype(np.float64) - ref[:, 0]) <= CDF_ATOL)
    assert np.array_equal((dev[:, 1] > 0), (ref[:, 1] > 0))
    np.testing.assert_allclose(dev[:, 1], ref[:, 1], rtol=PDF_RTOL, atol=0)
    assert np.array_equal(dev[:, 2].view(np.uint32), ref[:, 2].view(np.uint32))
    assert np.all(dev[:, 3] == 0)
    if ref[-1, 0] > 0:
        assert dev[-1, 0] == 1.0


def check_importance(dev, ref) -> None:
    assert (dev is None) == (ref is None)
    if ref is None:
        return
    (dm, dc, dp), (rm, rc, rp) = dev, ref
    assert dc.shape == rc.shape and dm.shape == rm.shape
    assert np.all(np.abs(dm.astype(np.float64) - rm) <= CDF_ATOL)
    assert np.all(np.abs(dc.astype(np.float64) - rc) <= CDF_ATOL)
    assert dm[-1] == 1.0 and np.all(dc[:, -1] == 1.0)
    assert np.all(np.diff(dm) >= 0) and np.all(np.diff(dc, axis=1) >= 0)
    np.testing.assert_allclose(dp, rp, rtol=PDF_RTOL, atol=0)


@pytest.fixture(scope="module")
def gpu():
    t = hx.PathTracer(0)
    yield t
    t.close()


@pytest.mark.parametrize("name,args", [("materials", (0,)), ("cornell", (0,)), ("instanced_small", (0,)),
                                       ("mixed", (3, 20000, 256))])
def test_emissive_table_matches_oracle(gpu, name, args):
    scene = hx.ProceduralScene(name, *args)
    gpu.set_scene(scene)
    dev = gpu.emissive_table(scene.triangle_count)
    check_emissive(dev, oracle.OracleScene(scene.view).emissive_table())


def emissive_heightfield(grid: int = 1024):
    """The heightfield with its ground material made emissive: one light
    table entry per triangle (2,093,058 at grid 1024), 512 scan blocks."""
    scene = hx.ProceduralScene("heightfield", 0, grid, 64)
    view = copy_view(scene.view)
    mats = np.frombuffer(ctypes.string_at(view.materials, 368 * view.material_count), dtype=np.uint8).copy()
    f = mats.view(np.float32).reshape(-1, 92)
    f[:, 4:7] = (0.9, 0.6, 0.3)   # emission colour
    f[:, 11] = 2.0                # emissionWeight
    view.materials = mats.ctypes.data
    view.material_emission_power = None
    return scene, view, mats


def test_emissive_table_two_million_triangles(gpu):
    scene, view, keep = emissive_heightfield()
    gpu.set_scene(view)
    dev = gpu.emissive_table(view.triangle_count)
    ref = oracle.OracleScene(view).emissive_table()
    assert np.count_nonzero(ref[:, 1]) == view.triangle_count
    check_emissive(dev, ref)
    del keep


@pytest.mark.parametrize("slot_kind", ["plain", "udim"])
def test_material_power_from_records_matches_oracle(gpu, slot_kind):
    scene = hx.ProceduralScene("materials", 0)
    u = scene.arrays()["materials"].view(np.uint32).reshape(-1, 92)
    slots = [int(x) for x in u[:, 68:92].reshape(-1) if x != MISSING]
    slot = next(x for x in slots if bool(x & 0x80000000) == (slot_kind == "udim"))
    view, keep, _ = with_emission_texture(scene, slot)
    gpu.set_scene(view)
    check_emissive(gpu.emissive_table(view.triangle_count), oracle.OracleScene(view).emissive_table())
    del keep


@pytest.mark.parametrize("name,args", [("heightfield", (0, 16, 2048)), ("materials", (0,)),
                                       ("mixed_lights", (1, 5000, 256))])
def test_dome_importance_matches_oracle(gpu, name, args):
    scene = hx.ProceduralScene(name, *args)
    gpu.set_scene(scene)
    o = oracle.OracleScene(scene.view)
    domes = 0
    for light in range(scene.view.light_count):
        ref = o.dome_importance(light)
        check_importance(gpu.dome_importance(light), ref)
        domes += ref is not None
    assert domes >= 1


def dark_dome_view(scene):
    """The scene's view with its dome texture replaced by zeros."""
    view = copy_view(scene.view)
    lights = np.frombuffer(ctypes.string_at(view.lights, 224 * max(view.light_count, 1)), dtype=np.uint32)
    ti = int(lights[52])
    tex = (hx.TextureView * view.texture_count)()
    for i in range(view.texture_count):
        tex[i] = view.textures[i]
    zeros = np.zeros(tex[ti].width * tex[ti].height * 4, dtype=np.float32)
    tex[ti].format = 2
    tex[ti].texels = zeros.ctypes.data
    view.textures = tex
    return view, (tex, zeros)


def test_dark_dome_is_sampled_uniformly(gpu):
    scene = hx.ProceduralScene("mixed", 5, 3000, 128)
    view, keep = dark_dome_view(scene)
    gpu.set_scene(view)
    o = oracle.OracleScene(view)
    assert gpu.dome_importance(0) is None and o.dome_importance(0) is None
    cam = scene.camera(48, 48)
    img = gpu.render(cam, 48, 48, 0, 8, 16)
    ref = o.render(cam, 48, 48, 0, 8, 16, precision=64, threads=cpu_threads()) / 16
    err = rel_mse(img, ref)
    print(f"dark dome, mixed 3000 tris: rel-MSE vs FP64 {err:.3e}")
    assert err < 1e-4
    del keep


def host_tables_view(scene):
    """The reference's packing of the same scene: the emissive table given in
    the view and the dome's importance slot naming an RGBA32F texture (r
    conditional CDF, g marginal CDF, b pdf) appended after the others; both
    are the oracle's restatement of the host build."""
    o = oracle.OracleScene(scene.view)
    view = copy_view(scene.view)
    emissive = np.ascontiguousarray(o.emissive_table())
    view.emissive = emissive.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    lights = np.frombuffer(ctypes.string_at(view.lights, 224 * max(view.light_count, 1)), dtype=np.uint32).copy()
    marginal, cond, pdf = o.dome_importance(0)
    h, w = pdf.shape
    imp = np.ones((h, w, 4), dtype=np.float32)
    imp[..., 0] = cond
    imp[..., 1] = marginal[:, None]
    imp[..., 2] = pdf
    tex = (hx.TextureView * (view.texture_count + 1))()
    for i in range(view.texture_count):
        tex[i] = view.textures[i]
    tex[view.texture_count] = hx.TextureView(w, h, 2, 0, imp.ctypes.data)
    lights.reshape(-1, 56)[0, 55] = view.texture_count
    view.textures = tex
    view.texture_count += 1
    view.lights = lights.ctypes.data
    return view, (emissive, lights, imp, tex), (marginal, cond, pdf)


def test_host_built_tables_are_used_unchanged(gpu):
    scene = hx.ProceduralScene("materials", 0)
    view, keep, tables = host_tables_view(scene)
    gpu.set_scene(view)
    assert np.array_equal(gpu.emissive_table(view.triangle_count), keep[0])
    dev = gpu.dome_importance(0)
    for a, b in zip(dev, tables):
        assert np.array_equal(a, b)
    cam = scene.camera(48, 48)
    img = gpu.render(cam, 48, 48, 0, 8, 16)
    ref = oracle.OracleScene(view).render(cam, 48, 48, 0, 8, 16, precision=64, threads=cpu_threads()) / 16
    err = rel_mse(img, ref)
    print(f"host-built tables, materials: rel-MSE vs FP64 {err:.3e}")
    assert err < 1e-4
    del keep


def test_emissive_heightfield_frame_matches_oracle(gpu):
    """Mesh-light sampling from the 2M-entry device table: every triangle
    is a light (next-event estimation picks them by the device CDF)."""
    scene, view, keep = emissive_heightfield(256)
    gpu.set_scene(view)
    cam = scene.camera(96, 54)
    img = gpu.render(cam, 96, 54, 0, 8, 16)
    ref = oracle.OracleScene(view).render(cam, 96, 54, 0, 8, 16, precision=64, threads=cpu_threads()) / 16
    err = rel_mse(img, ref)
    print(f"emissive heightfield 256: rel-MSE vs FP64 {err:.3e}")
    assert err < 1e-4
    del keep
