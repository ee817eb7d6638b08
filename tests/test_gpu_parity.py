"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (from the integrator's contract):
  * closest-hit primitive ids bit-exact, t/u/v bit-exact with the FP32
    oracle (0 ulp); against the FP64 evaluation of the same triangle, t is
    within 32 eps x scene extent and within 2 ulp for >= 95% of hits;
  * images rel-MSE < 1e-4 against the FP64 oracle on identical seeds;
  * progressive batching identical to sequential single-sample renders
    within 1e-5 (reference tests/path_tracer_tests.cpp:143-150).
"""

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import cutout_view, rel_mse, scene_rays, ulp_distance

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _check_hits(scene, rays, gpu, ref, t64):
    """Bit-exact against the FP32 oracle; against FP64 the watertight test's
    error is absolute (its shear cancels for origins near the plane), so it
    is bounded by 32 eps x the largest vertex-to-origin extent and reported
    as the fraction within 2 ulp."""
    assert np.array_equal(gpu["primitive"], ref["primitive"]), (
        f"{np.count_nonzero(gpu['primitive'] != ref['primitive'])} primitive mismatches")
    hit = ref["primitive"] != MISS
    assert np.array_equal(gpu["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
    assert np.array_equal(gpu["u"][hit].view(np.uint32), ref["u"][hit].view(np.uint32))
    assert np.array_equal(gpu["v"][hit].view(np.uint32), ref["v"][hit].view(np.uint32))
    ok = hit & np.isfinite(t64)
    a = scene.arrays()
    tri = a["positions"][:, :3].astype(np.float64)[a["indices"][gpu["primitive"][ok]]]
    scale = np.abs(tri - rays[ok, None, 0:3].astype(np.float64)).max(axis=(1, 2))
    err = np.abs(gpu["t"][ok].astype(np.float64) - t64[ok])
    # Edge functions and T carry absolute errors of a few eps x extent; the
    # division by det (proportional to |cos| of the ray/normal angle)
    # amplifies them for grazing rays.
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cos = np.abs((nrm * rays[ok, 4:7].astype(np.float64)).sum(axis=1))
    assert np.all(err * cos <= 16 * 2.0 ** -24 * scale)
    within = float(np.mean(ulp_distance(gpu["t"][ok], t64[ok].astype(np.float32)) <= 2))
    print(f"hits {int(hit.sum())}: t within 2 ulp of FP64 for {within:.4f}; max |dt|/scale/eps "
          f"{float((err / scale).max(initial=0) / 2.0 ** -24):.2f}")
    assert within >= 0.95
    return int(hit.sum())


@pytest.mark.parametrize("seed", [0, 1])
def test_closest_hit_cornell_bit_exact(tracer, seed):
    scene = hx.ProceduralScene("cornell", seed)
    tracer.set_scene(scene)
    cam = scene.camera(256, 256)
    rays = scene_rays(cam, 20000, seed, inside=(np.array([-1.0, 0.0, -1.0]), np.array([1.0, 2.0, 1.0])))
    gpu = tracer.trace_closest(rays)
    ref, t64 = oracle.OracleScene(scene.view).trace_closest(rays, brute_force=True)
    assert _check_hits(scene, rays, gpu, ref, t64) > 10000


def test_closest_hit_heightfield_bit_exact(tracer):
    scene = hx.ProceduralScene("heightfield", 3, 256, 64)
    tracer.set_scene(scene)
    cam = scene.camera(320, 180)
    rays = scene_rays(cam, 50000, 7, inside=(np.array([-50.0, -5.0, -50.0]), np.array([50.0, 30.0, 50.0])))
    gpu = tracer.trace_closest(rays)
    o = oracle.OracleScene(scene.view)
    ref, t64 = o.trace_closest(rays)
    brute, _ = o.trace_closest(rays[:2000], brute_force=True)
    assert np.array_equal(brute["primitive"], ref["primitive"][:2000])
    assert _check_hits(scene, rays, gpu, ref, t64) > 20000


def test_occluded_matches_oracle(tracer):
    scene = hx.ProceduralScene("materials", 0)
    tracer.set_scene(scene)
    cam = scene.camera(320, 160)
    rays = scene_rays(cam, 20000, 11, inside=(np.array([-5.0, 0.0, -3.0]), np.array([5.0, 4.0, 4.0])))
    rays[:, 7] = np.random.default_rng(5).uniform(0.01, 12.0, len(rays)).astype(np.float32)
    gpu = tracer.trace_occluded(rays)
    ref = oracle.OracleScene(scene.view).trace_occluded(rays)
    assert np.array_equal(gpu, ref)


def _render_pair(tracer, scene, w, h, spp, bounces=8):
    tracer.set_scene(scene)
    cam = scene.camera(w, h)
    gpu = tracer.render(cam, w, h, 0, bounces, spp)
    ref = oracle.OracleScene(scene.view).render(cam, w, h, 0, bounces, spp, precision=64) / spp
    return gpu, ref


def test_render_cornell_matches_fp64_oracle(tracer):
    gpu, ref = _render_pair(tracer, hx.ProceduralScene("cornell", 0), 64, 64, 16)
    assert gpu.shape == (64, 64, 4)
    assert np.all(gpu[..., 3] == 1.0)
    err = rel_mse(gpu, ref)
    print("cornell rel-MSE vs FP64 oracle", err)
    assert err < 1e-4


def test_render_materials_matches_fp64_oracle(tracer):
    gpu, ref = _render_pair(tracer, hx.ProceduralScene("materials", 0), 96, 48, 8)
    err = rel_mse(gpu, ref)
    print("materials rel-MSE vs FP64 oracle", err)
    assert err < 1e-4


def test_render_heightfield_matches_fp64_oracle(tracer):
    gpu, ref = _render_pair(tracer, hx.ProceduralScene("heightfield", 1, 128, 256), 80, 45, 4)
    err = rel_mse(gpu, ref)
    print("heightfield rel-MSE vs FP64 oracle", err)
    assert err < 1e-4


@pytest.mark.parametrize("kind,args,size", [("materials", (0,), (96, 48, 8)), ("mixed", (3, 20000, 64), (80, 48, 4))],
                         ids=["materials", "mixed"])
def test_render_cutout_matches_fp64_oracle(tracer, kind, args, size):
    """Scenes with cut-outs: every other material at opacity 0.5, so closest
    hits go through the any-hit loop (one random number per candidate in
    (t, primitive) order, :1228-1243) and shadow rays confirm occluders of
    opacity >= 0.5 (:855-889)."""
    scene = hx.ProceduralScene(kind, *args)
    view, keep = cutout_view(scene)
    w, h, spp = size
    tracer.set_scene(view)
    cam = scene.camera(w, h)
    gpu = tracer.render(cam, w, h, 0, 8, spp)
    ref = oracle.OracleScene(view).render(cam, w, h, 0, 8, spp, precision=64) / spp
    err = rel_mse(gpu, ref)
    print(kind, "cut-out rel-MSE vs FP64 oracle", err)
    assert err < 1e-4
    del keep


def test_sample_offset_gives_the_streams_of_later_sample_numbers(tracer):
    """hdcx_set_sample_offset seeds sample n as n + offset: the frame equals
    the oracle's samples [offset, offset + spp), and differs from offset 0."""
    scene = hx.ProceduralScene("cornell", 0)
    tracer.set_scene(scene)
    cam = scene.camera(48, 48)
    base = tracer.render(cam, 48, 48, 0, 8, 8)
    tracer.set_sample_offset(5000)
    try:
        shifted = tracer.render(cam, 48, 48, 0, 8, 8)
    finally:
        tracer.set_sample_offset(0)
    ref = oracle.OracleScene(scene.view).render(cam, 48, 48, 5000, 8, 8, precision=64) / 8
    assert rel_mse(shifted, ref) < 1e-4
    assert not np.array_equal(shifted, base)


def test_progressive_batching_matches_sequential(tracer):
    # tests/path_tracer_tests.cpp:143-150 on the reference's own scene.
    scene = hx.ProceduralScene("triangle", 0)
    tracer.set_scene(scene)
    cam = scene.camera(64, 64)
    tracer.render(cam, 64, 64, 0)
    seq = tracer.render(cam, 64, 64, 1)
    batched = tracer.render(cam, 64, 64, 0, 8, 2)
    assert batched.shape == seq.shape
    assert np.max(np.abs(batched - seq)) < 1e-5


def test_reference_triangle_scene_inequalities(tracer):
    # Qualitative checks of tests/path_tracer_tests.cpp:155-166.
    scene = hx.ProceduralScene("triangle", 0)
    tracer.set_scene(scene)
    assert tracer.has_geometry()
    cam = scene.camera(64, 64)
    px = tracer.render(cam, 64, 64, 0, 8, 64)
    assert px.shape == (64, 64, 4)
    assert np.all(np.isfinite(px))
    center = px[32, 32, :3].sum()
    corner = px[0, 0, :3].sum()
    assert abs(center - corner) > 0.01
    assert px[32, 32, 1] > px[32, 32, 0]  # green texture sampled
    assert np.all(px[..., 3] == 1.0)
    assert tracer.render(cam, 0, 64, 0).size == 0


def test_empty_scene_has_no_geometry(tracer):
    tracer.set_scene(None)
    assert not tracer.has_geometry()
    cam = hx.Camera.make([0, 0, 0], [-1, -1, -1], [2, 0, 0], [0, 2, 0])
    assert tracer.render(cam, 16, 16, 0).size == 0
