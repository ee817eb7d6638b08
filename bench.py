"""Benchmark of the spectral path-tracing hot path on one MI355X.

Workload (one step = one progressive frame): the seeded heightfield scene of
a 1024 x 1024 vertex grid (2,093,058 triangles, 8-octave fBm), GGX dielectric
over Burley materials, a 2048 x 1024 HDR lat-long dome with luminance
importance sampling; 1920 x 1080, 16 spp (x 3 wavelength lanes), 8 bounces
with Russian roulette; FP32 transport, FP64 film.

Printed: one JSON line with Mpaths/s (1 path = 1 pixel sample = 3 lanes),
the roofline of the dominant kernel from HIP-event timing of every launch in
the timed region plus node/triangle counts from a counting pass, the parity
rel-MSE of the frame against the FP64 CPU oracle on a pixel subset with the
same seeds, and the CPU baseline (the FP32 restatement timed on the host's
cores over a bounded sample of the same frame).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mpaths/s + rel-MSE at equal time, radiance-cache vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_QUERY = 2 * (111 * 64 + 4 * 64 * 64 + 64 * 3)  # 47,360 algorithmic FLOP per inference
WIDTH, HEIGHT, SPP, BOUNCES = 1920, 1080, 16, 8
GRID, DOME = 1024, 2048


TRAFFIC = ROOT / "profiles" / "hbm_traffic_heightfield1024_1080p_16spp.md"


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        for line in TRAFFIC.read_text().splitlines():
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return float(parts[2])
    except OSError:
        pass
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats: dict) -> dict:
    """Algorithmic bytes per launch-type (no cache credit).

    extend: per ray 32 B ray in + 16 B hit out + nodes x 128 B + triangles x 48 B.
    shadow: per ray 32 B ray + 4 B contribution in/out + nodes x 128 + triangles x 48.
    """
    out = {}
    er = stats["extend_rays_counted"]
    if er:
        per = 32 + 16 + stats["nodes_visited"] / er * stats["node_bytes"] + \
            stats["triangles_tested"] / er * stats["triangle_bytes"]
        out["extend"] = {"bytes_per_ray": per, "nodes_per_ray": stats["nodes_visited"] / er,
                         "triangles_per_ray": stats["triangles_tested"] / er}
    sr = stats["shadow_rays_counted"]
    if sr:
        per = 32 + 8 + stats["shadow_nodes_visited"] / sr * stats["node_bytes"] + \
            stats["shadow_triangles_tested"] / sr * stats["triangle_bytes"]
        out["shadow"] = {"bytes_per_ray": per, "nodes_per_ray": stats["shadow_nodes_visited"] / sr,
                         "triangles_per_ray": stats["shadow_triangles_tested"] / sr}
    return out


# Bytes the miss kernel moves per path vertex (every vertex of the bounce):
MISS_VERTEX_READ = 20     # queue entry 4, hit 16
MISS_READ = 40            # on a miss: flags 4, MIS pdfs 8, throughput 4, radiance 4, direction 16, wavelength 4
MISS_WRITE = 4            # on a miss: radiance (environment added)
MISS_HIT_WRITE = 4        # on a hit: hit-queue entry
# Bytes the shade kernel moves per surface hit (single-level scene, the
# simple-material variant that the heightfield runs, no radiance cache),
# from its loads and stores:
SHADE_VERTEX_READ = 80    # hit-queue entry 4, hit 16, ray origin 16 + direction 16, rng 4, wavelength 4,
                          # throughput 4, radiance 4, flags 4, MIS pdfs 8 (no medium in this variant)
SHADE_HIT_READ = 112      # the triangle's shading record: world positions 3 x 12, corner normals 3 x 12,
                          # UVs 3 x 8, material id 4, padded to 7 x 16 (one 128-B line per triangle)
SHADE_HIT_WRITE = 12      # rng, throughput, radiance
SHADE_KEPT_WRITE = 56     # next ray origin + direction 32, flags 4, MIS pdfs 8, survivor-queue entry 4;
                          # lane pairing: extend-queue entry 4 + follower mask 4 (counted for every kept path)
SHADE_SHADOW_WRITE = 88   # shadow slot: origin + tmax 16, direction + contribution 16, light 4, list end 4,
                          # shadow-queue entry 4; lane pairing: the slot read back 36, leader entry + mask 8


def miss_bytes(timed: dict) -> dict:
    """Algorithmic bytes of the miss launches of the timed steps (path state
    of every vertex; the dome texels a miss reads are a scene-wide table in
    L2/MALL and are not counted)."""
    items = timed["miss"]["items"]
    hits = timed["shade"]["items"]
    misses = max(items - hits, 0)
    total = items * MISS_VERTEX_READ + misses * (MISS_READ + MISS_WRITE) + hits * MISS_HIT_WRITE
    return {"bytes": total, "bytes_per_vertex": total / max(items, 1), "vertices": items,
            "hit_fraction": hits / max(items, 1)}


def shade_bytes(timed: dict) -> dict:
    """Algorithmic bytes of the shade launches of the timed steps: the bytes
    unique to each surface hit (its path state, hit record, triangle
    attributes and the state, queue and shadow-ray records it writes).
    Scene-wide tables that every vertex reads (material records, light
    records, dome texels and importance tables, textures: a few MB in
    L2/MALL) are not counted per vertex."""
    hits = timed["shade"]["items"]
    kept = max(timed["miss"]["items"] - timed["raygen"]["items"], 0)  # every later vertex was kept by the one before
    queued = timed["shadow_queued_paths"]
    total = (hits * (SHADE_VERTEX_READ + SHADE_HIT_READ + SHADE_HIT_WRITE)
             + kept * SHADE_KEPT_WRITE + queued * SHADE_SHADOW_WRITE)
    return {"bytes": total, "bytes_per_hit": total / max(hits, 1), "hits": hits, "shadow_queued_paths": queued}


def cpu_share() -> dict:
    """Host CPUs this process may use: the scheduler affinity, capped by a
    cgroup CPU quota and by OMP_NUM_THREADS when either is set (the share of
    the host a job is granted), with the host's model and logical CPU count."""
    import oracle

    info = oracle.host_cpu()
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = affinity
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
            share = min(share, max(1, int(quota)))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return {"threads": share, "model": info["model"], "logical_cpus": info["logical_cpus"],
            "affinity": affinity, "cgroup_quota": quota, "omp_num_threads": omp}


def cpu_baseline(scene, cam, share: dict, target_s: float = 25.0) -> dict:
    """FP32 CPU restatement (the -O3 build for this host's x86-64 level, SAH
    BVH) timed on the host's CPU share over the heightfield frame: a pilot on
    every 60th row picks the row stride so that the timed sample takes at
    most about target_s seconds (every row when that fits: about 14 s on a
    16-thread share), all columns, all 16 spp."""
    import oracle

    threads = share["threads"]
    o = oracle.OracleScene(scene.view)
    pilot = np.arange(0, HEIGHT, 60, dtype=np.uint32)
    t0 = time.perf_counter()
    o.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, precision=32, rows=pilot, col_step=1, threads=threads)
    rate = len(pilot) * WIDTH * SPP / (time.perf_counter() - t0)
    step = 1
    while step < 60 and (HEIGHT // step) * WIDTH * SPP / rate > target_s:
        step += 1
    rows = np.arange(0, HEIGHT, step, dtype=np.uint32)
    t0 = time.perf_counter()
    o.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, precision=32, rows=rows, col_step=1, threads=threads)
    dt = time.perf_counter() - t0
    paths = len(rows) * WIDTH * SPP
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"FP32 restatement (oracle/oracle.cpp, {oracle.BUILD_USED}, binned-SAH BVH), rows 0,{step},.."
                      f" x {WIDTH} columns x {SPP} spp = {paths} paths in {dt:.2f} s on {threads} threads",
            "host": {k: share[k] for k in ("model", "logical_cpus", "affinity", "cgroup_quota", "omp_num_threads")}}


def parity(scene, cam, frame: np.ndarray) -> dict:
    """rel-MSE of the GPU frame against the FP64 oracle, and against the FP32
    restatement, with the same seeds on every 16th row and 4th column."""
    import oracle

    rows = np.arange(0, HEIGHT, 16, dtype=np.uint32)
    o = oracle.OracleScene(scene.view)
    threads = cpu_share()["threads"]
    gpu = frame[rows.astype(np.int64)][:, ::4, :3].astype(np.float64)
    out = {"pixels": int(gpu.shape[0] * gpu.shape[1]), "sample": "rows 0,16,.. x columns 0,4,.. at the bench's seeds"}
    for precision, key in ((64, "rel_mse_vs_fp64_oracle"), (32, "rel_mse_vs_fp32_restatement")):
        ref = o.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, precision=precision, rows=rows, col_step=4,
                       threads=threads) / SPP
        out[key] = float(np.mean((gpu - ref) ** 2 / (ref * ref + 0.01)))
    return out


REFERENCES = {"4096spp": ROOT / "tests" / "golden" / "c1_reference_4096spp.npz",
              "32768spp": ROOT / "tests" / "golden" / "c1_reference_32768spp.npz"}


def score(frame: np.ndarray, ref) -> dict:
    """rel-MSE of a frame against an FP64 oracle reference on its pixel
    subset; "debiased" subtracts the reference's own noise (estimated from
    its two independent halves), an unbiased estimate of the error against
    the converged image."""
    rows, cols = ref["rows"].astype(np.int64), ref["cols"].astype(np.int64)
    y = 0.5 * (ref["half_a"] + ref["half_b"])
    floor = float(np.mean((ref["half_a"] - ref["half_b"]) ** 2 / (y * y + 0.01))) / 4.0
    x = frame[rows][:, cols, :3].astype(np.float64)
    e = float(np.mean((x - y) ** 2 / (y * y + 0.01)))
    return {"rel_mse": e, "debiased": e - floor, "noise_floor": floor, "mean_ratio": float(x.mean() / y.mean())}


CACHE_CONFIG = dict(train_samples=65536, terminate_bounce=1, train_steps=4, learning_rate=1e-2, ema_decay=0.99)
WARM_W, WARM_H, WARM_SPP = 240, 135, 2  # training-only frames: the same view at 1/64 of the pixels


def _timed_render(tracer, cam, w, h, index, spp) -> float:
    t0 = time.perf_counter()
    tracer.render(cam, w, h, index, BOUNCES, spp, readback=False)  # synchronous
    return time.perf_counter() - t0


def cache_frame_model(hx, scene, cam, seed: int) -> tuple:
    """Time of one cached 1080p frame as fixed + per-spp cost (its training
    included), fitted on a separate tracer and network from 4- and 12-spp
    frames: only used to plan how many samples fit a budget."""
    t = hx.PathTracer(0)
    try:
        t.set_bvh_options(hx.BVH_HOST_SAH, 2048, True)
        t.set_scene(scene)
        t.set_radiance_cache(True, seed=seed, **CACHE_CONFIG)
        for i in range(3):
            _timed_render(t, cam, WIDTH, HEIGHT, 500 + i, 4)
        times = {n: min(_timed_render(t, cam, WIDTH, HEIGHT, 0, n) for _ in range(2)) for n in (4, 12)}
    finally:
        t.close()
    per_spp = max((times[12] - times[4]) / 8, 1e-6)
    return max(times[4] - 4 * per_spp, 0.0), per_spp


def cold_start_session(tracer, cam, budget_s: float, model: tuple, seed: int, warm_fraction: float,
                       max_warm_frames: int = 200) -> dict:
    """The radiance-cache method from a fresh network within a wall-clock
    budget: training-only frames (the same view at 240 x 135, 2 spp) for
    `warm_fraction` of the budget, then 1080p frames accumulated
    progressively with the cache (each trains the network too), the last one
    given the samples that still fit. Everything after the network's
    initialisation is inside the budget."""
    fixed, per_spp = model
    tracer.set_radiance_cache(True, seed=seed, **CACHE_CONFIG)  # a new seed re-initialises the network
    t0 = time.perf_counter()
    warm = 0
    while warm < max_warm_frames and time.perf_counter() - t0 < warm_fraction * budget_s:
        tracer.render(cam, WARM_W, WARM_H, 2 * warm, BOUNCES, WARM_SPP, readback=False)
        warm += 1
    warm_s = time.perf_counter() - t0
    spp_done, frames = 0, 0
    while True:
        left = budget_s - (time.perf_counter() - t0)
        spp = min(16, int((0.97 * left - fixed) / per_spp))
        if spp < 1:
            break
        tracer.render(cam, WIDTH, HEIGHT, spp_done, BOUNCES, spp, readback=False)
        spp_done += spp
        frames += 1
    total_s = time.perf_counter() - t0
    return {"warm_frames": warm, "warm_s": warm_s, "frames": frames, "spp": spp_done, "seconds": total_s}


def equal_time(hx, scene, tracer, cam, plain_frame: np.ndarray, plain_ms: float, sessions: int = 8) -> dict:
    """Radiance cache vs plain at equal wall time on the heightfield
    workload, from a cold start, scored against FP64 oracle references
    (4096 spp on every 8th row and column: 32,400 pixels; 32,768 spp on
    every 32nd). For budgets of K plain frames (16 spp each, accumulated
    progressively), the cache method starts from an untrained network and
    spends the same wall time on training-only frames and cached frames.
    rel-MSE is heavy-tailed (a rare bright path shared by both methods can
    dominate an image), so each budget is run as `sessions` independent
    sessions (disjoint sample streams, a fresh network each) and the ratio
    of the mean errors is reported. Also: a warm network (trained on 50
    earlier frames, outside the budget) rendering one frame in one plain
    frame's time."""
    refs = {k: np.load(p) for k, p in REFERENCES.items() if p.exists()}
    model = cache_frame_model(hx, scene, cam, 77)
    out = {"config": {**CACHE_CONFIG, "terminate_vertex": "second (index 1)"},
           "frame_model": {"fixed_ms": model[0] * 1e3, "per_spp_ms": model[1] * 1e3},
           "warm_frame": f"{WARM_W}x{WARM_H}, {WARM_SPP} spp, 30% of the budget", "sessions": sessions, "budgets": []}
    for k in (1, 4):
        per = {key: {"plain": [], "cache": []} for key in refs}
        runs = []
        for j in range(sessions):
            tracer.set_sample_offset(1_000_000 + 10_000 * j)
            tracer.set_radiance_cache(False)
            t0 = time.perf_counter()
            for f in range(k):
                tracer.render(cam, WIDTH, HEIGHT, 16 * f, BOUNCES, 16, readback=False)
            plain_s = time.perf_counter() - t0
            plain = tracer.read_film(WIDTH, HEIGHT)
            session = cold_start_session(tracer, cam, plain_s, model, 1000 + 100 * k + j, 0.3)
            if session["spp"] == 0:
                raise RuntimeError("the cold-start session rendered no cached frame within its budget")
            cached = tracer.read_film(WIDTH, HEIGHT)
            session["plain_s"] = plain_s
            runs.append(session)
            for key, ref in refs.items():
                per[key]["plain"].append(score(plain, ref)["rel_mse"])
                per[key]["cache"].append(score(cached, ref)["rel_mse"])
        tracer.set_sample_offset(0)
        entry = {"plain_frames": k, "plain_spp": 16 * k, "budget_ms": 1e3 * float(np.mean([r["plain_s"] for r in runs])),
                 "cache_spp": [r["spp"] for r in runs], "cache_warm_frames": [r["warm_frames"] for r in runs],
                 "cache_ms": [1e3 * r["seconds"] for r in runs]}
        for key in refs:
            p, c = np.array(per[key]["plain"]), np.array(per[key]["cache"])
            entry[key] = {"plain_rel_mse": p.tolist(), "cache_rel_mse": c.tolist(),
                          "mean_plain": float(p.mean()), "mean_cache": float(c.mean()),
                          "ratio_of_means": float(p.mean() / c.mean()),
                          "median_ratio": float(np.median(p / c)), "sessions_cache_better": int((c < p).sum())}
        out["budgets"].append(entry)
        log(f"equal time, {k} plain frame(s) ~ {entry['budget_ms']:.0f} ms, {sessions} cold sessions: "
            + ", ".join(f"{key} ratio of means {entry[key]['ratio_of_means']:.2f}" for key in refs))
    out["warm_network_one_frame"] = warm_network_frame(hx, scene, tracer, cam, plain_frame, plain_ms, refs, model)
    head = out["budgets"][0]
    out["headline"] = {"budget": "one plain frame (16 spp), cold start, mean over sessions", "reference": "4096spp",
                       "ratio_plain_over_cache": head["4096spp"]["ratio_of_means"] if "4096spp" in head else None}
    return out


def warm_network_frame(hx, scene, tracer, cam, plain_frame, plain_ms: float, refs: dict, model) -> dict:
    """A network trained on 50 earlier frames of 2 spp (outside the budget)
    renders one cached frame in one plain frame's time (the interactive
    case); scored against the references when given."""
    tracer.set_radiance_cache(True, seed=1, **CACHE_CONFIG)
    for i in range(50):
        tracer.render(cam, WARM_W, WARM_H, 1000 + 2 * i, BOUNCES, WARM_SPP, readback=False)
    fixed, per_spp = model
    spp = max(1, int((0.98 * plain_ms * 1e-3 - fixed) / per_spp))
    tracer.reset_kernel_stats()
    tracer.set_instrumentation(1)
    before = tracer.rc_stats()
    t0 = time.perf_counter()
    tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, spp, readback=False)
    cache_ms = (time.perf_counter() - t0) * 1e3
    k = tracer.kernel_stats()
    tracer.set_instrumentation(0)
    frame = tracer.read_film(WIDTH, HEIGHT)
    stats = tracer.rc_stats()
    queries = stats["queries"] - before["queries"]
    tracer.set_radiance_cache(False)
    warm = {"spp": spp, "ms": cache_ms, "plain_ms": plain_ms, "train_loss": stats["last_loss"],
            "queries": int(queries), "train_ms": stats["train_ms"] - before["train_ms"]}
    for key, ref in refs.items():
        p, c = score(plain_frame, ref), score(frame, ref)
        warm[key] = {"plain": p, "cache": c, "ratio_plain_over_cache": p["rel_mse"] / max(c["rel_mse"], 1e-30)}
    infer_ms = k["cache"]["ms"]
    warm["mlp_infer_ms"] = infer_ms
    if infer_ms > 0 and queries > 0:
        tflops = queries * MLP_FLOP_PER_QUERY / (infer_ms * 1e-3) / 1e12
        warm["mlp_mfma"] = {"bound": "mfma", "achieved": tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": tflops / BF16_PEAK_TFLOPS, "flop_per_query": MLP_FLOP_PER_QUERY}
    return warm


def batch_workload(hx, threads: int, frames: int = 1, graph: bool = True, cache: bool = True) -> dict:
    """64 independent seeded mixed-material scenes (triangle targets
    log-uniform in [1e3, 5e5], square images of 128..1024 pixels, 8..64 spp,
    8 bounces) rendered together by the batch renderer: autotuned wave /
    tile / grid sizes, per-wave hipGraph replay. Reports aggregate Mpaths/s
    and each scene's rel-MSE against the FP64 oracle on 32 pixels (four rows
    of eight) at the same seeds. With `cache`, also per-scene radiance caches
    against the plain batch at equal wall time, both scored per scene
    against FP64 oracle references of the same pixels (2 x 256 spp)."""
    import oracle

    rng = np.random.default_rng(2024)
    specs = []
    for i in range(64):
        tris = int(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        res = int(rng.choice([128, 256, 384, 512, 640, 768, 896, 1024]))
        spp = int(rng.integers(8, 65))
        specs.append((i, tris, res, spp))
    b = hx.BatchRenderer(0)
    scenes = []
    t0 = time.perf_counter()
    for i, tris, res, spp in specs:
        scene = hx.ProceduralScene("mixed", i, tris, 256)
        scenes.append(scene)
        b.add_scene(scene, hx.BVH_DEVICE_LBVH, 2048)
        b.set_view(i, scene.camera(res, res), res, res, spp)
    setup_s = time.perf_counter() - t0
    log(f"batch: 64 scenes set up in {setup_s:.1f} s")
    t0 = time.perf_counter()
    if graph:
        tuning = b.autotune(BOUNCES, None)
    else:
        # Plain launches (for profilers that do not follow graph launches),
        # with the sizes the autotuner picks on this workload.
        b.set_tuning(3 << 21, 16384, 4096, False)
        tuning = b.tuning()
    tune_s = time.perf_counter() - t0
    b.render(0, BOUNCES)  # warm-up frame
    times = []
    for f in range(frames):
        t0 = time.perf_counter()
        b.render(0, BOUNCES)
        times.append(time.perf_counter() - t0)
    stats = b.stats()
    log(f"batch: plain frame {min(times) * 1e3:.0f} ms")
    plain_films = [b.read_film(i) for i in range(len(specs))]
    errs, refs, subsets = [], [], []
    ref_half = 256
    for (i, tris, res, spp), scene in zip(specs, scenes):
        rows = np.linspace(res // 8, res - res // 8, 4).astype(np.uint32)
        col_step = res // 8
        subsets.append((rows.astype(np.int64), col_step))
        o = oracle.OracleScene(scene.view)
        cam = scene.camera(res, res)
        ref = o.render(cam, res, res, 0, BOUNCES, spp, precision=64, rows=rows, col_step=col_step, threads=threads) / spp
        x = plain_films[i][rows.astype(np.int64)][:, ::col_step, :3].astype(np.float64)
        errs.append(float(np.mean((x - ref) ** 2 / (ref * ref + 0.01))))
        if i % 16 == 15:
            log(f"batch: oracle scenes {i + 1}/64")
        if cache:
            halves = [o.render(cam, res, res, 100000 + h * ref_half, BOUNCES, ref_half, precision=64, rows=rows,
                               col_step=col_step, threads=threads) / ref_half for h in range(2)]
            refs.append(halves)
    ms = min(times) * 1e3
    out = {"workload": "64 mixed-material scenes, 1e3..5e5 triangles (log-uniform), 128^2..1024^2, 8..64 spp, "
                       "8 bounces, one batch (seeded)",
           "scenes": 64, "triangles": int(sum(sc.triangle_count for sc in scenes)),
           "pixel_samples": stats["pixel_samples"], "ms_per_frame": ms,
           "mpaths_per_s": stats["pixel_samples"] / (ms * 1e-3) / 1e6, "waves": stats["waves"],
           "graph_launches": stats["graph_launches"], "tuning": tuning, "autotune_s": tune_s, "setup_s": setup_s,
           "rel_mse_vs_fp64_oracle": {"max": max(errs), "median": float(np.median(errs)), "per_scene": errs,
                                      "sample": "32 pixels per scene (4 rows x 8 columns), same seeds"}}
    if cache:
        out["equal_time"] = batch_equal_time(b, specs, scenes, plain_films, ms, refs, subsets)
    b.close()
    return out


def batch_equal_time(b, specs, scenes, plain_films, plain_ms: float, refs, subsets, warm_frames: int = 20) -> dict:
    """Per-scene radiance caches vs the plain batch at equal wall time. The
    caches are warm: warm_frames renders of 2 spp per scene train them (not
    timed); then every view's spp is scaled so that one render (its training
    included) takes about the plain render's time."""
    def score(film, halves, sub):
        rows, step = sub
        a, c = halves
        y = 0.5 * (a + c)
        floor = float(np.mean((a - c) ** 2 / (y * y + 0.01))) / 4.0
        x = film[rows][:, ::step, :3].astype(np.float64)
        e = float(np.mean((x - y) ** 2 / (y * y + 0.01)))
        return e, e - floor

    def set_spp(scale: float):
        out = []
        for (i, _, res, spp), scene in zip(specs, scenes):
            n = max(1, int(round(spp * scale)))
            b.set_view(i, scene.camera(res, res), res, res, n)
            out.append(n)
        return out

    config = dict(train_samples=16384, terminate_bounce=1, train_steps=4, learning_rate=1e-2, seed=1, ema_decay=0.99)
    b.set_radiance_cache(True, **config)
    for (i, _, res, _), scene in zip(specs, scenes):
        b.set_view(i, scene.camera(res, res), res, res, 2)
    t0 = time.perf_counter()
    for f in range(warm_frames):
        b.render(1000 + 2 * f, BOUNCES)
    warm_s = time.perf_counter() - t0
    log(f"batch cache: {warm_frames} warm frames in {warm_s:.1f} s")
    set_spp(1.0)
    t0 = time.perf_counter()
    b.render(5000, BOUNCES)
    unit_s = time.perf_counter() - t0
    spps = set_spp(plain_ms * 1e-3 / unit_s)
    t0 = time.perf_counter()
    b.render(0, BOUNCES)
    cache_ms = (time.perf_counter() - t0) * 1e3
    log(f"batch cache: equal-time frame {cache_ms:.0f} ms")
    rc = b.rc_stats()
    plain, cached = [], []
    for i in range(len(specs)):
        plain.append(score(plain_films[i], refs[i], subsets[i]))
        cached.append(score(b.read_film(i), refs[i], subsets[i]))
    b.set_radiance_cache(False)
    for (i, _, res, spp), scene in zip(specs, scenes):
        b.set_view(i, scene.camera(res, res), res, res, spp)
    ratio = [p[0] / max(c[0], 1e-30) for p, c in zip(plain, cached)]
    return {"config": {**config, "train_samples_per_scene": config["train_samples"]}, "warm_frames": warm_frames,
            "warm_spp": 2, "warm_s": warm_s, "plain_ms": plain_ms, "cache_ms": cache_ms,
            "cache_spp_scale": spps[0] / specs[0][3], "cache_spp": spps,
            "plain_rel_mse": [p[0] for p in plain], "cache_rel_mse": [c[0] for c in cached],
            "plain_rel_mse_debiased": [p[1] for p in plain], "cache_rel_mse_debiased": [c[1] for c in cached],
            "ratio_plain_over_cache": ratio, "median_ratio": float(np.median(ratio)),
            "scenes_improved": int(sum(r > 1.0 for r in ratio)),
            "mean_plain_rel_mse": float(np.mean([p[0] for p in plain])),
            "mean_cache_rel_mse": float(np.mean([c[0] for c in cached])),
            "cache_queries": rc["queries"], "mean_loss": rc["mean_loss"],
            "reference": "FP64 oracle, same 32 pixels per scene, two independent halves of 256 spp"}


def cornell_workload(hx, threads: int, frames: int = 3) -> dict:
    """Seeded Cornell box (5 wall quads + 2 rotated boxes = 34 triangles, one
    0.25 x 0.25 rect light, Lambert albedos), 256 x 256, 64 spp, 8 bounces:
    GPU throughput, the frame's rel-MSE against the FP64 oracle over every
    pixel (same seeds), and the FP32 restatement timed over the whole frame
    on the host's cores."""
    import oracle

    w = h = 256
    spp = 64
    scene = hx.ProceduralScene("cornell", 0)
    cam = scene.camera(w, h)
    t = hx.PathTracer(0)
    try:
        t.set_scene(scene)
        t.render(cam, w, h, 0, BOUNCES, spp, readback=False)
        times = []
        for _ in range(frames):
            t0 = time.perf_counter()
            t.render(cam, w, h, 0, BOUNCES, spp, readback=False)
            times.append(time.perf_counter() - t0)
        frame = t.read_film(w, h)[..., :3].astype(np.float64)
    finally:
        t.close()
    o = oracle.OracleScene(scene.view)
    t0 = time.perf_counter()
    ref32 = o.render(cam, w, h, 0, BOUNCES, spp, precision=32, threads=threads) / spp
    cpu_s = time.perf_counter() - t0
    ref64 = o.render(cam, w, h, 0, BOUNCES, spp, precision=64, threads=threads) / spp
    paths = w * h * spp
    gpu_s = min(times)
    return {"workload": "Cornell box, 34 triangles, 256x256, 64 spp x 3 lanes, 8 bounces (seed 0)",
            "ms_per_frame": gpu_s * 1e3, "mpaths_per_s": paths / gpu_s / 1e6,
            "rel_mse_vs_fp64_oracle": float(np.mean((frame - ref64) ** 2 / (ref64 * ref64 + 0.01))),
            "rel_mse_vs_fp32_restatement": float(np.mean((frame - ref32) ** 2 / (ref32 * ref32 + 0.01))),
            "pixels_compared": w * h,
            "cpu_baseline": {"value": paths / cpu_s / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
                             "sample": f"FP32 restatement over the whole frame ({paths} paths) in {cpu_s:.2f} s"}}


def instanced_workload(hx, frames: int = 2) -> dict:
    """10,000 instances (seeded uniform scale, uniform rotation, positions in
    a slab) of a level-5 icosphere (20,480 triangles) over a ground quad:
    204.8 M instanced triangles, mixed materials per instance, 64 analytic
    lights (the 64-light cap drops the last in id order) plus a 2048 x 1024
    dome; 1920 x 1080, 32 spp, 8 bounces, FP64 film. Two-level BVH: host SAH
    per prototype, SAH instance tree over widened world boxes. Instanced hits
    and images are bit-identical to the flattened scene's (GPU tests at 40
    instances), so parity follows from the flattened oracle there."""
    w, h, spp = 1920, 1080, 32
    scene = hx.ProceduralScene("instanced", 0, 10000, DOME)
    cam = scene.camera(w, h)
    t = hx.PathTracer(0)
    try:
        t0 = time.perf_counter()
        t.set_scene(scene)
        set_scene_s = time.perf_counter() - t0
        info = t.bvh_info()
        t.render(cam, w, h, 0, BOUNCES, spp, readback=False)
        t.reset_kernel_stats()
        t.set_instrumentation(1)
        times = []
        for f in range(frames):
            t0 = time.perf_counter()
            t.render(cam, w, h, 0, BOUNCES, spp, readback=False)
            times.append(time.perf_counter() - t0)
        timed = t.kernel_stats()
        frame = t.read_film(w, h)
        # Counting pass: nodes (instance tree and prototype trees) and
        # triangles per ray.
        t.reset_kernel_stats()
        t.set_instrumentation(2)
        t.render(cam, w, h, 0, BOUNCES, spp, readback=False)
        counted = t.kernel_stats()
        t.set_instrumentation(0)
    finally:
        t.close()
    alg = algorithmic_bytes(counted)
    traversal = {}
    for k in ("extend", "shadow"):
        if k in alg and timed[k]["ms"] > 0:
            rays = timed[k]["items"]
            achieved = rays * alg[k]["bytes_per_ray"] / (timed[k]["ms"] * 1e-3) / 1e9
            traversal[k] = {**alg[k], "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS}
    ms = sum(times) / frames * 1e3
    paths = w * h * spp
    return {"workload": "10,000 instances of a 20,480-triangle icosphere (204.8 M instanced triangles) + ground, mixed "
                        "materials, 63 lights + dome, 1920x1080, 32 spp x 3 lanes, 8 bounces, two-level BVH (seeded)",
            "instances": scene.instance_count, "instanced_triangles": 10000 * 20480,
            "stored_triangles": scene.triangle_count, "ms_per_frame": ms, "mpaths_per_s": paths / (ms * 1e-3) / 1e6,
            "mrays_per_s": (timed["extend"]["items"] + timed["shadow"]["items"]) / (sum(times)) / 1e6,
            "kernels_ms_per_frame": {k: timed[k]["ms"] / frames for k in ("extend", "miss", "shade", "walk", "shadow")},
            "traversal": traversal,
            "set_scene_s": set_scene_s, "bvh": {k: info[k] for k in ("node_count", "max_stack", "depth", "build_ms")},
            "frame_mean_rgb": [float(x) for x in frame[..., :3].reshape(-1, 3).mean(axis=0)]}


def copy_calibration() -> int:
    """A plain copy of known size: 256 MiB read + 256 MiB written per launch.
    Its FETCH_SIZE / WRITE_SIZE readings give the scale that turns those
    counters into HBM bytes for the path-tracing kernels."""
    import torch

    n = 256 << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n, "copy_calibration_bytes_written": n, "launches": 4}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache equal-time comparison")
    ap.add_argument("--bvh", choices=("lbvh", "sah"), default="sah",
                    help="BVH builder: device LBVH with SAH top levels, or host binned SAH")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-scene batch workload (about a minute and a half)")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-scene batch workload")
    ap.add_argument("--no-instanced", action="store_true", help="skip the 10,000-instance two-level BVH workload")
    ap.add_argument("--instanced-only", action="store_true", help="only the 10,000-instance workload")
    ap.add_argument("--cache-frame-only", action="store_true",
                    help="only the radiance cache's warm-network frame at one plain frame's time (heightfield)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch workload with plain launches instead of graph replay, no autotuning")
    ap.add_argument("--no-batch-cache", action="store_true",
                    help="batch workload without the per-scene radiance caches at equal time")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy kernel 4 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.batch_only:
        import hdcodex_amd as hx
        import oracle

        oracle.use_fastest_build()
        print(json.dumps({"batch": batch_workload(hx, cpu_share()["threads"], graph=not args.no_graph,
                                                  cache=not args.no_batch_cache)}),
              flush=True)
        return 0
    if args.cache_frame_only:
        import hdcodex_amd as hx

        scene = hx.ProceduralScene("heightfield", 0, GRID, DOME)
        cam = scene.camera(WIDTH, HEIGHT)
        tracer = hx.PathTracer(0)
        tracer.set_bvh_options(hx.BVH_HOST_SAH, 2048, True)
        tracer.set_scene(scene)
        tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=False)
        t0 = time.perf_counter()
        tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=False)
        plain_ms = (time.perf_counter() - t0) * 1e3
        plain = tracer.read_film(WIDTH, HEIGHT)
        model = cache_frame_model(hx, scene, cam, 77)
        refs = {k: np.load(p) for k, p in REFERENCES.items() if p.exists()}
        print(json.dumps({"cache_frame": warm_network_frame(hx, scene, tracer, cam, plain, plain_ms, refs, model)}),
              flush=True)
        tracer.close()
        return 0
    if args.instanced_only:
        import hdcodex_amd as hx

        print(json.dumps({"instanced": instanced_workload(hx)}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import hdcodex_amd as hx
    import oracle

    oracle.use_fastest_build()  # the -O3 CPU build for this host (parity and baseline legs)
    log("building scene")
    scene = hx.ProceduralScene("heightfield", 0, GRID, DOME)
    cam = scene.camera(WIDTH, HEIGHT)
    tracer = hx.PathTracer(0)
    tracer.set_bvh_options(hx.BVH_DEVICE_LBVH if args.bvh == "lbvh" else hx.BVH_HOST_SAH, 2048, True)
    t0 = time.perf_counter()
    tracer.set_scene(scene)
    set_scene_s = time.perf_counter() - t0
    bvh = tracer.bvh_info()
    # Build time of the other builder on the same scene (reported, not timed in the steps).
    other = hx.PathTracer(0)
    other.set_bvh_options(hx.BVH_HOST_SAH if args.bvh == "lbvh" else hx.BVH_DEVICE_LBVH, 2048, True)
    other.set_scene(scene)
    other_bvh = other.bvh_info()
    other.close()
    # set_scene on fresh tracers, the process's one-time HIP set-up paid:
    # upload, validation, BVH build and the device light tables, with each
    # builder.
    set_scene_warm_s = {}
    for name, builder in (("device_lbvh", hx.BVH_DEVICE_LBVH), ("host_sah", hx.BVH_HOST_SAH)):
        again = hx.PathTracer(0)
        again.set_bvh_options(builder, 2048, True)
        t0 = time.perf_counter()
        again.set_scene(scene)
        set_scene_warm_s[name] = time.perf_counter() - t0
        again.close()
    log(f"set_scene {set_scene_s:.2f} s, {scene.triangle_count} triangles")

    for i in range(args.warmup):
        tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=False)
        log(f"warmup {i} done")
    tracer.reset_kernel_stats()
    tracer.set_instrumentation(1)  # HIP events around every launch of the timed steps
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=False)  # synchronous
        times.append(time.perf_counter() - t0)
        log(f"step {i}: {times[-1] * 1e3:.1f} ms")
    timed = tracer.kernel_stats()
    frame = tracer.read_film(WIDTH, HEIGHT)
    # The same frame with the RGBA32F result copied to host memory (the
    # boundary's output buffer), outside the timed steps.
    t0 = time.perf_counter()
    tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=True)
    pcie_s = time.perf_counter() - t0

    # Counting pass (separate instantiation of the traversal kernels).
    tracer.reset_kernel_stats()
    tracer.set_instrumentation(2)
    tracer.render(cam, WIDTH, HEIGHT, 0, BOUNCES, SPP, readback=False)
    counted = tracer.kernel_stats()
    tracer.set_instrumentation(0)
    alg = algorithmic_bytes(counted)

    total_s = float(sum(times))
    ms_per_step = total_s / args.steps * 1e3
    paths = WIDTH * HEIGHT * SPP
    mpaths = paths * args.steps / total_s / 1e6
    extend_rays = timed["extend"]["items"]
    shadow_rays = timed["shadow"]["items"]
    mrays = (extend_rays + shadow_rays) / total_s / 1e6

    kernels = {}
    for k in hx.KERNELS:
        kernels[k] = {"ms_per_step": timed[k]["ms"] / args.steps, "launches_per_step": timed[k]["launches"] / args.steps,
                      "items_per_step": timed[k]["items"] / args.steps}
        traffic = measured_traffic(k)
        if traffic and timed[k]["launches"]:
            # Measured HBM bytes per launch (rocprofv3 counters) over this run's average launch time.
            avg_s = timed[k]["ms"] * 1e-3 / timed[k]["launches"]
            kernels[k]["hbm_GBs"] = traffic / avg_s / 1e9
            kernels[k]["hbm_frac"] = kernels[k]["hbm_GBs"] / HBM_PEAK_GBS
    dominant = max(("extend", "miss", "shade", "shadow"), key=lambda k: timed[k]["ms"])
    # Rooflines of the four path kernels: algorithmic bytes of the timed
    # launches over their HIP-event time.
    roof = {}
    for k in ("extend", "shadow"):
        if k in alg and timed[k]["ms"] > 0:
            rays = extend_rays if k == "extend" else shadow_rays
            roof[k] = {**alg[k], "alg_bytes": rays * alg[k]["bytes_per_ray"], "rays_per_launch":
                       rays / max(timed[k]["launches"], 1)}
    if timed["shade"]["ms"] > 0:
        roof["shade"] = shade_bytes(timed)
        roof["shade"]["alg_bytes"] = roof["shade"].pop("bytes")
    if timed["miss"]["ms"] > 0:
        roof["miss"] = miss_bytes(timed)
        roof["miss"]["alg_bytes"] = roof["miss"].pop("bytes")
    for k, r in roof.items():
        launches = max(timed[k]["launches"], 1)
        r["achieved_GBs"] = r["alg_bytes"] / (timed[k]["ms"] * 1e-3) / 1e9
        r["frac"] = r["achieved_GBs"] / HBM_PEAK_GBS
        r["avg_launch_ms"] = timed[k]["ms"] / launches
        r["alg_bytes_per_launch"] = r["alg_bytes"] / launches
    roofline = None
    if dominant in roof:
        r = roof[dominant]
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": r["achieved_GBs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": r["frac"], "traffic": measured_traffic(dominant),
                    "alg_bytes_per_launch": r["alg_bytes_per_launch"], "avg_launch_ms": r["avg_launch_ms"],
                    "others": {k: {"frac": v["frac"], "achieved": v["achieved_GBs"],
                                   "alg_bytes_per_launch": v["alg_bytes_per_launch"],
                                   "avg_launch_ms": v["avg_launch_ms"], "traffic": measured_traffic(k)}
                               for k, v in roof.items() if k != dominant}}
        for k, v in roofline["others"].items():
            if v["frac"] > 1.0:
                # Node and triangle reads are algorithmic bytes whether L2 or
                # HBM serves them; the counter traffic is what HBM moved.
                v["note"] = ("algorithmic rate above the HBM peak: L2/MALL serve most of the node and triangle "
                             "reads; HBM traffic per launch is `traffic`")

    result = {
        "metric": METRIC,
        "value": mpaths,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"heightfield {GRID}x{GRID} grid ({scene.triangle_count} tris), {DOME}x{DOME // 2} HDR "
                               f"dome, {WIDTH}x{HEIGHT}, {SPP} spp x 3 lanes, {BOUNCES} bounces, plain path tracer",
                   "width": WIDTH, "height": HEIGHT, "spp": SPP, "bounces": BOUNCES, "radiance_cache": False},
        "mrays_per_s": mrays,
        "pcie_inclusive": {"ms_per_step": pcie_s * 1e3, "mpaths_per_s": paths / pcie_s / 1e6,
                           "note": "same frame with the 33 MB RGBA32F readback to host memory"},
        "lane_paths_per_s_M": 3 * mpaths,
        "set_scene_s": set_scene_s,
        "set_scene_warm_s": set_scene_warm_s,
        "bvh": {"builder": args.bvh, "build_ms": bvh["build_ms"], "nodes": bvh["node_count"],
                "max_stack": bvh["max_stack"], "depth": bvh["depth"], "top_clusters": bvh["top_clusters"],
                "other_builder": "sah" if args.bvh == "lbvh" else "lbvh", "other_build_ms": other_bvh["build_ms"],
                "other_nodes": other_bvh["node_count"]},
        "dominant_kernel": dominant,
        "kernels": kernels,
        "kernel_rooflines": roof,
        "roofline": roofline,
    }
    if not args.no_cache:
        log("radiance cache vs plain at equal time")
        result["equal_time"] = equal_time(hx, scene, tracer, cam, frame, ms_per_step)
    if not args.no_parity:
        log("parity vs FP64 oracle (subset)")
        result["parity"] = parity(scene, cam, frame)
        result["parity_rel_mse_vs_fp64_oracle"] = result["parity"]["rel_mse_vs_fp64_oracle"]
    log("cornell workload")
    result["cornell"] = cornell_workload(hx, cpu_share()["threads"])
    if not args.no_instanced:
        log("instanced workload")
        result["instanced"] = instanced_workload(hx)
    if not args.no_batch:
        log("batch workload")
        result["batch"] = batch_workload(hx, cpu_share()["threads"], graph=not args.no_graph,
                                         cache=not args.no_batch_cache)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(scene, cam, cpu_share())
    tracer.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
