"""Host-side checks without a GPU: the C ABI exports, the scene packer's
ordering and byte layouts (reference BuildScene,
src/gpu/vulkan_path_tracer.cpp:2604-2993), scene validation and failure
reporting, and oracle renders with closed-form answers."""

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import rel_mse, with_emission_texture

ROOT = Path(__file__).resolve().parent.parent
MISSING = 0xFFFFFFFF


def _header_symbols():
    text = (ROOT / "include" / "hdcx.h").read_text()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(hdcx_\w+)\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(hx.LIB_PATH))
    declared = _header_symbols()
    assert "hdcx_render" in declared and "hdcx_set_scene" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(hx.EXPORTED_SYMBOLS) == declared


def test_null_arguments_are_rejected_without_throwing():
    lib = hx.library()
    assert lib.hdcx_render(None, None, 8, 8, 0, 8, 1, None) == hx.ERR_INVALID_ARGUMENT
    assert lib.hdcx_last_error()
    assert lib.hdcx_has_geometry(None) == 0
    assert lib.hdcx_set_scene(None, None) == hx.ERR_INVALID_ARGUMENT


def test_create_reports_missing_device_or_succeeds():
    h = ctypes.c_void_p()
    code = hx.library().hdcx_create(0, ctypes.byref(h))
    if code == hx.OK:
        hx.library().hdcx_destroy(h)
    else:
        assert code == hx.ERR_DEVICE
        assert hx.library().hdcx_last_error()
    assert hx.library().hdcx_create(-1, ctypes.byref(h)) in (hx.ERR_DEVICE,)


def _materials(a):
    m = a["materials"]
    f = m.view(np.float32).reshape(len(m), 92)
    u = m.view(np.uint32).reshape(len(m), 92)
    return f, u


def test_default_material_record():
    # :2768-2800; index 0 is the default grey 0.8.
    f, u = _materials(hx.ProceduralScene("cornell", 0).arrays())
    assert np.allclose(f[0, 0:4], [0.8, 0.8, 0.8, 0.0])
    assert np.allclose(f[0, 4:8], [0, 0, 0, 0.5])
    assert np.allclose(f[0, 8:12], [0.0, 1.0, 1.5, 0.0])
    assert np.allclose(f[0, 40:44], [0.0, 0.0, 0.0, 20.0])
    assert np.allclose(f[0, 60:64], [1, 1, 1, -1])
    assert np.all(u[0, 68:92] == MISSING)


def test_cornell_packing_follows_id_order():
    s = hx.ProceduralScene("cornell", 0)
    a = s.arrays()
    assert s.triangle_count == 34
    # Meshes flatten in id order: box_short, box_tall, then the 5 walls
    # back, ceiling, floor, left, right; materials are numbered 1.. in id
    # order of their ids.
    tm = a["triangle_materials"]
    assert list(tm[:12]) == [1] * 12 and list(tm[12:24]) == [2] * 12
    assert list(tm[24:]) == [3, 3, 4, 4, 5, 5, 6, 6, 7, 7]
    f, _ = _materials(a)
    assert np.all(f[1:, 28] == 0.0)  # specularWeight 0: pure Lambert
    assert np.all((f[1:, 0:3] >= 0.1) & (f[1:, 0:3] <= 0.9))
    # One rect light: radiance scale 100, selection power lum x scale x area.
    L = a["lights"].view(np.float32).reshape(-1, 56)
    Lu = a["lights"].view(np.uint32).reshape(-1, 56)
    assert a["light_count"] == 1
    assert L[0, 3] == pytest.approx(100.0)
    assert Lu[0, 54] == 1  # type Rect
    assert L[0, 51] == pytest.approx(1.0 * 100.0 * 0.0625)
    assert L[0, 7] == pytest.approx(0.0625)  # controls.w = area
    # No emissive material: the device-built table (restated by the oracle)
    # is all zero, so no triangle is sampled as a light.
    assert a["emissive"] is None and np.all(a["material_emission_power"] == 0.0)
    assert np.all(oracle.OracleScene(hx.ProceduralScene("cornell", 0).view).emissive_table() == 0.0)


def test_texture_order_plain_then_udim_sets():
    # :2627-2654: plain textures in snapshot order, then each UDIM set in
    # set-id order with its tiles sorted; the set's index carries the flag,
    # the tile mask and the base slot.
    a = hx.ProceduralScene("triangle", 0).arrays()
    assert a["textures"] == [(1, 1, 1), (1, 1, 2), (1, 1, 0), (4, 1, 2), (1, 1, 1), (1, 1, 1)]
    _, u = _materials(a)
    assert u[1, 68] == 0  # baseColorTexture "green#srgb" -> slot 0
    m = hx.ProceduralScene("materials", 0).arrays()
    _, mu = _materials(m)
    udim = [x for x in mu[:, 68] if x != MISSING and x & 0x80000000]
    assert udim, "materials scene binds a UDIM set"
    x = int(udim[0])
    assert (x >> 8) & 0x7FFFFF == 0b11 and (x & 0xFF) < len(m["textures"])


def test_dome_importance_slot_names_the_dome_texture():
    # The packer leaves the importance tables to the device: the dome's
    # importance slot (textureInfo.w) names its own texture (textureInfo.x).
    s = hx.ProceduralScene("heightfield", 0, 16, 64)
    L = np.frombuffer(ctypes.string_at(s.view.lights, 224), dtype=np.uint32)
    assert L[54] == 0 and L[55] == L[52] != MISSING
    assert s.view.emissive is None or not s.view.emissive


def _row_solid_angles(h: int, w: int) -> np.ndarray:
    v0 = np.arange(h) / h
    return 2 * math.pi / w * np.abs(np.cos(math.pi * (1 - (v0 + 1 / h))) - np.cos(math.pi * (1 - v0)))


def test_dome_importance_tables_are_one_distribution():
    """The oracle's restatement of the importance build (:1998-2066) on the
    2048 x 1024 benchmark dome: the three planes describe ONE distribution.
    pdf x texel solid angle sums to 1; each marginal step is its row's pdf
    mass; each conditional step is the texel's share of its row; the pdf is
    proportional to the texel luminance recomputed here from the texture."""
    s = hx.ProceduralScene("heightfield", 0, 16, 2048)
    o = oracle.OracleScene(s.view)
    marginal, cond, pdf = o.dome_importance(0)
    h, w = pdf.shape
    assert (w, h) == (2048, 1024)
    solid = _row_solid_angles(h, w)
    mass = pdf.astype(np.float64) * solid[:, None]
    assert float(mass.sum()) == pytest.approx(1.0, rel=1e-6)
    assert marginal[-1] == 1.0 and np.all(cond[:, -1] == 1.0)
    assert np.all(np.diff(marginal) >= 0) and np.all(np.diff(cond, axis=1) >= 0)
    row_mass = mass.sum(axis=1)
    assert np.allclose(np.diff(np.concatenate([[0.0], marginal.astype(np.float64)])), row_mass, atol=2e-7)
    share = mass / np.where(row_mass > 0, row_mass, 1.0)[:, None]
    assert np.allclose(np.diff(np.concatenate([np.zeros((h, 1)), cond.astype(np.float64)], axis=1), axis=1), share,
                       atol=2e-7)
    v = s.view
    tex = np.ctypeslib.as_array(ctypes.cast(v.textures[int(np.frombuffer(ctypes.string_at(v.lights, 224),
                                                                          dtype=np.uint32)[52])].texels,
                                            ctypes.POINTER(ctypes.c_float)), shape=(h, w, 4))
    lum = np.clip(tex[..., :3].astype(np.float64), 0, None) @ np.array([0.2126, 0.7152, 0.0722])
    ratio = pdf / np.where(lum > 0, lum, 1.0)
    assert np.allclose(ratio[lum > 1e-3], ratio[lum > 1e-3].mean(), rtol=1e-5)


def test_emissive_table_is_the_area_power_distribution():
    """The oracle's restatement of the emissive-triangle build (:2944-2993):
    probability = area x material power / total, recomputed here from the
    positions and the packer's per-material powers; the cdf is the running
    sum of those probabilities with its last entry 1."""
    s = hx.ProceduralScene("materials", 0)
    a = s.arrays()
    e = oracle.OracleScene(s.view).emissive_table()
    p = a["positions"][:, :3].astype(np.float64)[a["indices"]]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    power = a["material_emission_power"][a["triangle_materials"]].astype(np.float64)
    weight = area.astype(np.float32).astype(np.float64) * power
    prob = weight / weight.sum()
    assert np.count_nonzero(prob) > 0 and np.count_nonzero(prob) < len(prob)
    assert np.allclose(e[:, 2], area, rtol=1e-6)
    assert np.allclose(e[:, 1].astype(np.float64) * e[:, 2], prob, rtol=1e-6, atol=1e-12)
    assert e[-1, 0] == 1.0 and np.allclose(e[:, 0], np.cumsum(prob), atol=2e-7)
    f, _ = _materials(a)
    assert np.all((e[:, 1] > 0) == (f[a["triangle_materials"], 11] * power > 0))


@pytest.mark.parametrize("slot_kind", ["color", "plain", "udim"])
def test_material_power_rule_from_records(slot_kind):
    """A view without per-material powers derives them from its records:
    max(emissionWeight, 0) x the luminance of the emission color, or the
    emission texture's mean texel luminance (a UDIM set: mean of its tiles'
    means), recomputed here with numpy from the texels."""
    s = hx.ProceduralScene("materials", 0)
    a = s.arrays()
    _, mu = _materials(a)
    if slot_kind == "color":
        view = hx.SceneView()
        ctypes.memmove(ctypes.byref(view), ctypes.byref(s.view), ctypes.sizeof(view))
        view.material_emission_power = None
        keep, power = None, a["material_emission_power"]
    else:
        slots = [int(x) for x in mu[:, 68:92].reshape(-1) if x != MISSING]
        slot = next(x for x in slots if bool(x & 0x80000000) == (slot_kind == "udim"))
        view, keep, power = with_emission_texture(s, slot)
    got = oracle.OracleScene(view).emissive_table()
    want_view = hx.SceneView()
    ctypes.memmove(ctypes.byref(want_view), ctypes.byref(view), ctypes.sizeof(view))
    p = np.ascontiguousarray(power, dtype=np.float32)
    want_view.material_emission_power = p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    want = oracle.OracleScene(want_view).emissive_table()
    assert np.any(want[:, 1] > 0)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    del keep


def test_validation_of_device_table_inputs():
    s = hx.ProceduralScene("cornell", 0)
    view = hx.SceneView()
    ctypes.memmove(ctypes.byref(view), ctypes.byref(s.view), ctypes.sizeof(view))
    view.light_triangle_count = view.triangle_count + 1
    with pytest.raises(hx.HdcxError) as e:
        hx.validate_scene(view)
    assert e.value.code == hx.ERR_INVALID_SCENE and "light_triangle_count" in str(e.value)
    view.light_triangle_count = view.triangle_count
    hx.validate_scene(view)


def _corrupted(scene, field, value, index=0):
    a = scene.arrays()
    keep = {k: np.ascontiguousarray(v) for k, v in a.items() if isinstance(v, np.ndarray)}
    view = hx.SceneView()
    ctypes.memmove(ctypes.byref(view), ctypes.byref(scene.view), ctypes.sizeof(view))
    arr = keep[field]
    arr.reshape(-1)[index] = value
    ptr = {"positions": ctypes.c_float, "indices": ctypes.c_uint32, "triangle_materials": ctypes.c_uint32}[field]
    setattr(view, field, arr.ctypes.data_as(ctypes.POINTER(ptr)))
    return view, keep


@pytest.mark.parametrize("field,value", [("indices", 10 ** 6), ("positions", float("nan")),
                                         ("triangle_materials", 99)])
def test_validation_rejects_corrupt_scenes(field, value):
    scene = hx.ProceduralScene("cornell", 0)
    hx.validate_scene(scene.view)
    view, _keep = _corrupted(scene, field, value)
    with pytest.raises(hx.HdcxError) as e:
        hx.validate_scene(view)
    assert e.value.code == hx.ERR_INVALID_SCENE


def test_oracle_furnace_closed_form():
    # Lambert 0.5 floor under a uniform white dome: every pixel's expectation
    # is 0.5 x the white point (1.0346, 0.9796, 1.0289).
    s = hx.ProceduralScene("furnace", 0)
    o = oracle.OracleScene(s.view)
    cam = s.camera(32, 32)
    img = o.render(cam, 32, 32, 0, 8, 64, precision=64) / 64
    mean = img.reshape(-1, 3).mean(axis=0)
    assert np.allclose(mean, 0.5 * np.array([1.0346, 0.9796, 1.0289]), rtol=0.01)


def test_oracle_progressive_sum_is_additive():
    s = hx.ProceduralScene("cornell", 0)
    o = oracle.OracleScene(s.view)
    cam = s.camera(24, 24)
    two = o.render(cam, 24, 24, 0, 8, 2, precision=64)
    a = o.render(cam, 24, 24, 0, 8, 1, precision=64)
    b = o.render(cam, 24, 24, 1, 8, 1, precision=64)
    assert np.allclose(two, a + b, rtol=1e-12, atol=1e-14)


def test_oracle_fp32_tracks_fp64():
    s = hx.ProceduralScene("materials", 0)
    o = oracle.OracleScene(s.view)
    cam = s.camera(64, 32)
    r32 = o.render(cam, 64, 32, 0, 8, 4, precision=32) / 4
    r64 = o.render(cam, 64, 32, 0, 8, 4, precision=64) / 4
    assert rel_mse(r32, r64) < 1e-4


def test_oracle_bvh_matches_brute_force():
    s = hx.ProceduralScene("materials", 0)
    o = oracle.OracleScene(s.view)
    rng = np.random.default_rng(3)
    org = rng.uniform([-5, 0.1, -2], [5, 4, 5], size=(3000, 3)).astype(np.float32)
    d = rng.normal(size=(3000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = hx.rays_array(org, d)
    a, _ = o.trace_closest(rays)
    b, _ = o.trace_closest(rays, brute_force=True)
    assert np.array_equal(a["primitive"], b["primitive"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
