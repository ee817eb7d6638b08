// Wavefront spectral path tracer for MI355X (gfx950).
//
// One chunk of lane paths (3 per pixel sample) flows through
//   raygen -> [extend -> miss -> shade -> shadow] x maxBounces -> film
// with every stage a grid-stride kernel over a device-side queue whose
// length the previous stage wrote; the host never reads a count back.
// This file holds raygen, extend, the lane pairing, shadow, film and the
// ray queries; miss, shade and walk are in shade.hip. The queues are
// compacted with a wave ballot + prefix popcount and one atomic per wave
// (or per block, dev_wavefront.h).
//
// FP contraction is off for the whole integrator: every instantiation of a
// stage (one scene or a scene table, one or two levels, walks inline or
// deferred) then rounds the same operations the same way, and the FP32
// ray/triangle arithmetic follows the oracle's (built with
// -ffp-contract=off) bit for bit.
#pragma clang fp contract(off)
#include "kernel_api.h"
#include "dev_wavefront.h"
#include "rcache.h"

#include <type_traits>


namespace hdcx {

// ---------------------------------------------------------------------------
// raygen: seed, hero wavelength, jittered pinhole ray (:1198-1224).
// ---------------------------------------------------------------------------
// One lane path: seed, hero wavelength, jittered pinhole ray and the
// initial path state. cam: origin, lowerLeft - origin, horizontal, vertical.
__device__ __forceinline__ void startPath(const PathBuffers& pb, const CacheParams& cp, std::uint32_t pid,
                                          std::uint32_t ps, std::uint32_t lane, std::uint32_t pixel,
                                          std::uint32_t sampleNumber, std::uint32_t width, std::uint32_t height,
                                          const float* cam)
{
    const std::uint32_t px = pixel % width, py = pixel / width;
    std::uint32_t rng = hashState(pixel ^ (sampleNumber * 0x9e3779b9U) ^ 0xa511e9b3U);
    float wl, wl0;
    {
#pragma clang fp contract(off)
        const float shift = __uint2float_rn(hashState(pixel ^ 0x68bc21ebU)) * (1.0f / 4294967296.0f);
        const float phase = (__uint2float_rn(sampleNumber & 31U) + 0.5f) / 32.0f;
        const float step = __uint2float_rn(lane) / 3.0f;
        const float f = shift + phase + step;
        wl = 380.0f + 400.0f * (f - floorf(f));
        const float f0 = shift + phase;
        wl0 = 380.0f + 400.0f * (f0 - floorf(f0));
    }
    pb.lambda0[pid] = wl0;
    if (cp.enabled) pb.spread[pid] = make_float2(0.0f, 0.0f);
    // (A single render empties every training record before raygen,
    // LaunchClearRecords; in a batch, slots are handed out per wave by
    // raygenBatchKernel and the gather of the previous wave left every
    // handed-out record empty.)
    const float jx = randomFloat(rng), jy = randomFloat(rng);
    const float u = (static_cast<float>(px) + jx) / static_cast<float>(width);
    const float v = (static_cast<float>(py) + jy) / static_cast<float>(height);
    // vulkan_path_tracer.cpp:1211-1213 with lowerLeft - origin taken from
    // the host (rounded once from FP64): same direction, without the FP32
    // cancellation when |origin| is large.
    const V3 o = v3(cam[0], cam[1], cam[2]);
    const V3 d = normalize(v3(cam[3], cam[4], cam[5]) + v3(cam[6], cam[7], cam[8]) * u + v3(cam[9], cam[10], cam[11]) * v);
    pb.rayO[pid] = make_float4(o.x, o.y, o.z, 0.0f);
    pb.rayD[pid] = make_float4(d.x, d.y, d.z, 0.0f);
    pb.rng[pid] = rng;
    pb.wavelength[pid] = wl;
    pb.throughput[pid] = 1.0f;
    pb.radiance[pid] = 0.0f;
    // kPrevDelta: the MIS pdfs are read only after a vertex has set them;
    // the medium only with kInMedium, set together with it; shadow slots
    // only after shade has queued them. None of the three is written here.
    pb.flags[pid] = kEnvOnMiss | kPrevDelta;
    pb.queue[0][pid] = pid;
}

__global__ __launch_bounds__(kBlock) void raygenKernel(PathBuffers pb, ChunkDesc c, std::uint32_t pathCount,
                                                      CacheParams cp)
{
    const std::uint32_t pixelsInChunk = c.p1 - c.p0;
    for (std::uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x; pid < pathCount; pid += gridDim.x * blockDim.x) {
        const std::uint32_t ps = pid / 3U, lane = pid - ps * 3U;
        const std::uint32_t sOff = ps / pixelsInChunk;
        const std::uint32_t pixel = c.p0 + (ps - sOff * pixelsInChunk);
        startPath(pb, cp, pid, ps, lane, pixel, c.sampleBase + c.s0 + sOff, c.width, c.height, c.cam);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.counters[0] = pathCount;
        if (cp.enabled) *cp.queryCount = 0U;
    }
}

// Last tile of [a, b) whose offset (selected by `field`) is <= x.
template <class Field>
__device__ __forceinline__ std::uint32_t findTile(const BatchTile* tiles, std::uint32_t a, std::uint32_t b,
                                                  std::uint32_t x, Field field)
{
    while (b - a > 1U) {
        const std::uint32_t m = (a + b) >> 1;
        if (field(tiles[m]) <= x) a = m;
        else b = m;
    }
    return a;
}

// Batch raygen: the lane paths of the current wave (read from the device
// cursor, so that the same launch serves every wave of a captured graph).
// With the radiance cache on, a pixel sample trains when its index within
// the scene's frame (sample-major) plus the scene's offset is a multiple of
// the scene's stride; each of its lane paths takes a record slot.
__global__ __launch_bounds__(kBlock) void raygenBatchKernel(PathBuffers pb, BatchWaves bw, CacheParams cp)
{
    const std::uint32_t w = *bw.cursor;
    const std::uint32_t ta = bw.waveTiles[w], tb = bw.waveTiles[w + 1];
    const std::uint32_t paths = bw.wavePaths[w];
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (paths + stride - 1U) / stride * stride;  // every lane reaches the ballot
    for (std::uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x; pid < padded; pid += stride) {
        bool train = false;
        if (pid < paths) {
            const BatchTile& t = bw.tiles[findTile(bw.tiles, ta, tb, pid, [](const BatchTile& x) { return x.pathOffset; })];
            const BatchView& v = bw.views[t.scene];
            const std::uint32_t local = (pid - t.pathOffset) / 3U, lane = pid - t.pathOffset - local * 3U;
            const std::uint32_t tilePixels = t.p1 - t.p0;
            const std::uint32_t so = local / tilePixels;
            const std::uint32_t pixel = t.p0 + (local - so * tilePixels);
            pb.pathScene[pid] = t.scene;
            startPath(pb, cp, pid, local, lane, pixel, v.sampleBase + so, v.width, v.height, v.cam);
            if (cp.enabled) {
                const uint2 tr = cp.sceneTrain[t.scene];
                const std::uint32_t index = so * (v.width * v.height) + pixel;
                train = tr.x != 0U && (index + tr.y) % tr.x == 0U;
            }
        }
        if (cp.enabled) {
            const std::uint32_t slot = waveSlot(train, cp.recordCount);
            if (pid < paths) cp.trainSlot[pid] = train && slot < cp.recordSlots ? slot : 0xffffffffU;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) pb.counters[0] = paths;
}

// Path-vertex hit of a non-opaque scene: candidates in increasing
// (t, primitive) order, each confirmed with randomFloat <= opacity
// (the reference's any-hit loop :1233-1243 with a defined order).
template <bool kCount, bool k2L>
__device__ bool traceStochastic(const DeviceScene& s, const RayPre& r, std::uint32_t& rng, HitRec& hit,
                                TraversalStack& st, TraversalCounters* ctr)
{
    float loT = 0.001f;
    std::uint32_t loPrim = 0xffffffffU;
    for (;;) {
        HitRec h;
        if (!traceClosestS<kCount, k2L>(s, r, 0.001f, 10000.0f, loT, loPrim, h, st, ctr)) return false;
        if (randomFloat(rng) <= materialOpacity<k2L>(s, h.prim, h.u, h.v)) {
            hit = h;
            return true;
        }
        loT = h.t;
        loPrim = h.prim;
    }
}

// ---------------------------------------------------------------------------
// extend: closest hit for every queued path (:1228-1243).
//
// sharing: kShareNone traces every queued path. kShareCamera (the launch
// after raygen): the queue is every lane path of the launch in order,
// three per pixel sample, and the three lanes of a sample start with the
// same camera ray and the same random state (the seed and the jitter do
// not depend on the lane; only the wavelength does), so the ray is traced
// once, for lane 0. kShareLanes: the previous shade queued the paths that trace
// (pb.extendQueue) with a mask of the later lanes of their sample whose
// next ray it found equal (pb.follow). The hit (and, in a scene with
// cut-outs, the random state after the opacity draws) is written to the
// path and to the lanes that share its ray: the same values each lane
// would compute for itself.
// ---------------------------------------------------------------------------
// kOpaque: every scene of the launch is opaque (the host knows), so the
// cut-out candidate loop is compiled out and its registers with it.
// Occupancy, measured on the 10,000-instance frame: with speculative
// traversal the two-level kernel runs best held to 6 waves per SIMD
// (extend 517 ms per frame at 7 waves, 453 at 6, 466 at 5). One-level
// opaque scenes use extendRefillKernel; the generic one-level kernel is held
// to 6 waves (80 VGPRs).
template <bool kCount, class Scenes, bool k2L, bool kOpaque>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void extendKernel(Scenes scenes, PathBuffers pb, std::uint32_t in, int sharing,
                                                      TraversalCounters* ctr)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.counters[(in ^ 1U) * kCounterStride] = 0U;
        pb.counters[2 * kCounterStride] = 0U;
        pb.counters[3 * kCounterStride] = 0U;
        pb.counters[(4U + (in ^ 1U)) * kCounterStride] = 0U;
        pb.counters[6 * kCounterStride] = 0U;
        pb.counters[kHitCounter * kCounterStride] = 0U;
    }
    const std::uint32_t count = pb.counters[in * kCounterStride];
    const bool camera = sharing == kShareCamera && count % 3U == 0U;
    const bool leaders = sharing == kShareLanes;
    const std::uint32_t rays = camera ? count / 3U : leaders ? pb.counters[(4U + in) * kCounterStride] : count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.totals[0] += count;
        pb.totals[3] += rays;
    }
    const std::uint32_t* queue = leaders ? pb.extendQueue[in] : pb.queue[in];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rays; i += gridDim.x * blockDim.x) {
        const std::uint32_t pid = camera ? 3U * i : queue[i];
        const std::uint32_t followers = camera ? 3U : leaders ? pb.follow[pid] : 0U;
        const DeviceScene& s = scenes.Get(pid);
        const float4 o = pb.rayO[pid], d = pb.rayD[pid];
        const RayPre r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
        HitRec h;
        bool found;
        std::uint32_t rng = 0U;
        if (kOpaque || s.opaque) {
            found = traceClosestS<kCount, k2L>(s, r, 0.001f, 10000.0f, 0.0f, 0xffffffffU, h, st, ctr);
        } else {
            rng = pb.rng[pid];
            found = traceStochastic<kCount, k2L>(s, r, rng, h, st, ctr);
        }
        const float4 hit = make_float4(h.t, __uint_as_float(found ? h.prim : 0xffffffffU), h.u, h.v);
        if constexpr (kCount) {
            if (ctr && found) atomicAdd(&ctr->hits, 1ULL + static_cast<unsigned long long>(__popc(followers & 3U)));
        }
        pb.hit[pid] = hit;
        if (!kOpaque && !s.opaque) pb.rng[pid] = rng;
        for (std::uint32_t k = 0; k < 2U; ++k) {
            if (!(followers & (1U << k))) continue;
            pb.hit[pid + 1U + k] = hit;
            if (!kOpaque && !s.opaque) pb.rng[pid + 1U + k] = rng;
        }
    }
}

// ---------------------------------------------------------------------------
// extend with lane refill: one-level opaque scenes (the same hits as
// extendKernel). In extendKernel a wave traces 64 rays and takes the next
// 64 only when its slowest lane is done. Here a lane that has finished its
// ray takes the next one as soon as kRefillLanes lanes of the wave are idle:
// the wave walks the same 64-ray groups as the grid-stride loop (group
// g = wave + j * waves), lane-refilled in group order, so a wave's rays stay
// neighbours and every wave samples the whole queue. Each step is one
// speculative traversal round (closestRound). On the 1080p heightfield
// frame: refill alone left extend at 15.8 ms per frame, speculation alone
// gave 14.7 ms, both 13.3 ms (22 instead of 17 of 64 lanes active per VALU
// instruction, 22% fewer memory instructions). Held to 5 waves per SIMD (no
// spills): 13.2 ms, against 13.4 at 6, 13.9 at 7 and 16.0 at 8.
// ---------------------------------------------------------------------------
constexpr int kRefillLanes = 16;

template <bool kCount, class Scenes>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void extendRefillKernel(
    Scenes scenes, PathBuffers pb, std::uint32_t in, int sharing, TraversalCounters* ctr)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.counters[(in ^ 1U) * kCounterStride] = 0U;
        pb.counters[2 * kCounterStride] = 0U;
        pb.counters[3 * kCounterStride] = 0U;
        pb.counters[(4U + (in ^ 1U)) * kCounterStride] = 0U;
        pb.counters[6 * kCounterStride] = 0U;
        pb.counters[kHitCounter * kCounterStride] = 0U;
    }
    const std::uint32_t count = pb.counters[in * kCounterStride];
    const bool camera = sharing == kShareCamera && count % 3U == 0U;
    const bool leaders = sharing == kShareLanes;
    const std::uint32_t rays = camera ? count / 3U : leaders ? pb.counters[(4U + in) * kCounterStride] : count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.totals[0] += count;
        pb.totals[3] += rays;
    }
    const std::uint32_t* queue = leaders ? pb.extendQueue[in] : pb.queue[in];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    const std::uint64_t waves = static_cast<std::uint64_t>(gridDim.x) * (kBlock / 64);
    const std::uint64_t wave = static_cast<std::uint64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const unsigned long long below = (1ULL << laneId()) - 1ULL;
    // Rays handed out to this wave so far, counted in its own group order:
    // the v-th is ray (wave + (v / 64) * waves) * 64 + v % 64.
    std::uint64_t taken = 0;
    bool active = false;
    std::uint32_t pid = 0, followers = 0;
    const DeviceScene* sc = nullptr;  // the active ray's scene (set at refill)
    ClosestState c;
    c.exhausted = true;
    RayPre r = prepareRay(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f);
    unsigned long long nodes = 0, tris = 0;
    for (;;) {
        const unsigned long long lanes = __ballot(true);
        const unsigned long long idle = __ballot(!active);
        if (__popcll(idle) >= kRefillLanes || idle == lanes) {
            const std::uint64_t first = (wave + (taken >> 6) * waves) * 64U + (taken & 63U);
            if (idle == lanes && first >= rays) break;
            if (!active) {
                const std::uint64_t v = taken + static_cast<std::uint64_t>(__popcll(idle & below));
                const std::uint64_t k = (wave + (v >> 6) * waves) * 64U + (v & 63U);
                if (k < rays) {
                    const std::uint32_t i = static_cast<std::uint32_t>(k);
                    pid = camera ? 3U * i : queue[i];
                    followers = camera ? 3U : leaders ? pb.follow[pid] : 0U;
                    sc = &scenes.Get(pid);
                    startClosest(c, 10000.0f, st);
                    if (sc->nodeCount != 0) {
                        const float4 o = pb.rayO[pid], d = pb.rayD[pid];
                        r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
                        active = true;
                    } else {
                        // Empty scene: the ray misses.
                        const HitRec h = closestResult(c);
                        const float4 miss = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
                        pb.hit[pid] = miss;
                        for (std::uint32_t j = 0; j < 2U; ++j)
                            if (followers & (1U << j)) pb.hit[pid + 1U + j] = miss;
                    }
                }
            }
            taken += static_cast<std::uint64_t>(__popcll(idle));
        }
        if (!active) continue;
        // One scene: its arrays stay in scalar registers (per-lane pointers
        // would cost a 64-bit address computation per load).
        if constexpr (std::is_same<Scenes, SingleScene>::value)
            closestRound<kCount>(scenes.Get(0), r, 0.001f, 0.0f, 0xffffffffU, c, st, nodes, tris);
        else
            closestRound<kCount>(*sc, r, 0.001f, 0.0f, 0xffffffffU, c, st, nodes, tris);
        if (c.exhausted) {
            const HitRec h = closestResult(c);
            const float4 hit = make_float4(h.t, __uint_as_float(c.found ? h.prim : 0xffffffffU), h.u, h.v);
            if constexpr (kCount) {
                if (ctr && c.found) atomicAdd(&ctr->hits, 1ULL + static_cast<unsigned long long>(__popc(followers & 3U)));
            }
            pb.hit[pid] = hit;
            for (std::uint32_t k = 0; k < 2U; ++k)
                if (followers & (1U << k)) pb.hit[pid + 1U + k] = hit;
            active = false;
        }
    }
    if (kCount && ctr) {
        atomicAdd(&ctr->nodes, nodes);
        atomicAdd(&ctr->triangles, tris);
    }
}

// ---------------------------------------------------------------------------
// shadow: visibility of each queued path's shadow rays, added to its
// radiance in slot order (shadowVisibility :855-889, meshLightVisible
// :891-912). kOpaque as in extendKernel. Held to 6 waves per SIMD: on the
// 1080p heightfield frame 7.75 ms against 7.90 at 7 waves and 8.8 at 8
// (5 waves: the same as 6).
// ---------------------------------------------------------------------------
template <bool kCount, class Scenes, bool k2L, bool kOpaque>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void shadowKernel(Scenes scenes, PathBuffers pb, bool shared, TraversalCounters* ctr)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    __shared__ unsigned int tracedInBlock;
    if (threadIdx.x == 0) tracedInBlock = 0U;
    __syncthreads();
    unsigned int traced = 0U;
    const std::uint32_t queued = pb.counters[2 * kCounterStride];
    if (blockIdx.x == 0 && threadIdx.x == 0) pb.totals[1] += queued;
    // shared: shade grouped the lanes of a pixel sample with
    // equal shadow rays; each group's rays are traced once and the
    // visibility applied to every lane of the group.
    const std::uint32_t count = shared ? pb.counters[6 * kCounterStride] : queued;
    const std::uint32_t* queue = shared ? pb.shadowLeaders : pb.shadowQueue;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const std::uint32_t pid = queue[i];
        const DeviceScene& s = scenes.Get(pid);
        // Trace every slot first, keeping only the outcomes: the lanes'
        // wavelengths, radiance and the light controls are read afterwards,
        // so that they are not held in registers across the traversals.
        // (Per-slot values in named registers: an array indexed by the
        // rolled slot loop would live in scratch memory.)
        static_assert(kShadowSlots == 3, "three occluder distances");
        std::uint32_t occludedSlots = 0U;
        float occluder0 = 0.0f, occluder1 = 0.0f, occluder2 = 0.0f;
        int slots = 0;
        for (int k = 0; k < kShadowSlots; ++k) {
            const std::size_t idx = static_cast<std::size_t>(k) * pb.capacity + pid;
            const std::uint32_t light = pb.shadowLight[idx];
            if (light == 0xffffffffU) break;
            slots = k + 1;
            const float4 o = pb.shadowO[idx], d = pb.shadowD[idx];
            float tmax = o.w;
            const bool mesh = light == kShadowMeshLight;
            bool falloff = false;
            if (!mesh) {
                const F4 shadow = s.lights[light].shadow;
                if (shadow.x < 0.5f) continue;
                if (shadow.y >= 0.0f) tmax = fminf(tmax, shadow.y);
                falloff = shadow.y >= 0.0f && shadow.z > 0.0f;
            }
            if (tmax <= 0.001f) continue;
            ++traced;
            const RayPre r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
            bool occluded;
            float hitDistance = 0.0f;
            const bool opaque = kOpaque || s.opaque;
            if (opaque && !falloff) {
                occluded = traceAnyS<kCount, k2L>(s, r, 0.001f, tmax, st, ctr);
            } else {
                // Closest confirmed occluder: opacity >= 0.5 in a
                // non-opaque scene; its distance drives the falloff.
                float loT = 0.001f;
                std::uint32_t loPrim = 0xffffffffU;
                occluded = false;
                for (;;) {
                    HitRec h;
                    if (!traceClosestS<kCount, k2L>(s, r, 0.001f, tmax, loT, loPrim, h, st, ctr)) break;
                    if (opaque || materialOpacity<k2L>(s, h.prim, h.u, h.v) >= 0.5f) {
                        occluded = true;
                        hitDistance = h.t;
                        break;
                    }
                    loT = h.t;
                    loPrim = h.prim;
                }
            }
            if (occluded) occludedSlots |= 1U << k;
            if (k == 0) occluder0 = hitDistance;
            else if (k == 1) occluder1 = hitDistance;
            else occluder2 = hitDistance;
        }
        if (slots == 0) continue;
        const std::uint32_t followers = shared ? pb.shadowFollow[pid] : 0U;
        std::uint32_t member[3] = {pid, pid, pid};
        int members = 1;
        for (std::uint32_t k = 0; k < 2U; ++k)
            if (followers & (1U << k)) member[members++] = pid + 1U + k;
        float wls[3], rad[3];
        for (int m = 0; m < members; ++m) {
            wls[m] = pb.wavelength[member[m]];
            rad[m] = pb.radiance[member[m]];
        }
        for (int k = 0; k < slots; ++k) {
            const std::size_t idx = static_cast<std::size_t>(k) * pb.capacity + pid;
            const std::uint32_t light = pb.shadowLight[idx];
            const bool mesh = light == kShadowMeshLight;
            float4 shadow = make_float4(1.0f, -1.0f, -1.0f, 1.0f);
            float4 shadowColor = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            bool trace = true;
            if (!mesh) {
                const GpuLightRecord& L = s.lights[light];
                shadow = make_float4(L.shadow.x, L.shadow.y, L.shadow.z, L.shadow.w);
                shadowColor = make_float4(L.shadowColor.x, L.shadowColor.y, L.shadowColor.z, 0.0f);
                if (shadow.x < 0.5f) trace = false;
            }
            const bool occluded = trace && (occludedSlots >> k & 1U) != 0U;
            const bool falloffSet = !mesh && shadow.y >= 0.0f && shadow.z > 0.0f;
            // Visibility of each lane of the group with its own wavelength, and
            // its own contribution, in slot order.
            for (int m = 0; m < members; ++m) {
                const float wl = wls[m];
                const float contribution = pb.shadowD[static_cast<std::size_t>(k) * pb.capacity + member[m]].w;
                float visibility = 1.0f;
                if (occluded) {
                    if (mesh) {
                        visibility = 0.0f;
                    } else {
                        float strength = 1.0f;
                        if (falloffSet) {
                            const float f = fminf(shadow.z, fmaxf(shadow.y, 0.0f));
                            const float occluder = k == 0 ? occluder0 : k == 1 ? occluder1 : occluder2;
                            strength = clampf((shadow.y - occluder) / fmaxf(f, 1e-6f), 0.0f, 1.0f);
                            strength = powf(strength, fmaxf(shadow.w, 0.0f));
                        }
                        const V3 vis = v3(mixf(1.0f, shadowColor.x, strength), mixf(1.0f, shadowColor.y, strength),
                                          mixf(1.0f, shadowColor.z, strength));
                        visibility = reflectanceSpectrum(vis, wl);
                    }
                } else if (!mesh) {
                    visibility = reflectanceSpectrum(v3(1.0f), wl);
                }
                if (mesh) {
                    if (visibility > 0.0f) rad[m] += contribution;
                } else {
                    rad[m] += contribution * visibility;
                }
            }
        }
        for (int m = 0; m < members; ++m) pb.radiance[member[m]] = rad[m];
    }
    atomicAdd(&tracedInBlock, traced);
    __syncthreads();
    if (threadIdx.x == 0 && tracedInBlock) atomicAdd(&pb.totals[2], static_cast<unsigned long long>(tracedInBlock));
}

// Training targets of a chunk: for each recorded vertex j of a training
// path, the radiance the path gathered after vertex j's emission divided by
// the throughput that reached j (an unbiased estimate of the radiance
// scattered at j toward the previous vertex).
// Training sample of record i: the record's query with the target radiance
// after its vertex (the path's final radiance minus the radiance up to the
// vertex, over the throughput reaching it); false for an empty record or a
// non-finite target.
__device__ __forceinline__ bool trainingSample(const PathBuffers& pb, const CacheParams& cp, std::uint32_t i, RcQuery& r)
{
    r = cp.records[i];
    if (r.lane > 2U) return false;
    const float t = cp.recordThroughput[i];
    if (!(t > 0.0f)) return false;
    const float target = (pb.radiance[r.path] - r.target) / t;
    if (!isfinite(target)) return false;
    r.target = fmaxf(target, 0.0f);
    return true;
}

__global__ void gatherTrainingKernel(PathBuffers pb, CacheParams cp, std::uint32_t records)
{
    if (cp.recordCount) records = min(*cp.recordCount, cp.recordSlots) * cp.recordVertices;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < records; i += gridDim.x * blockDim.x) {
        RcQuery r;
        if (!trainingSample(pb, cp, i, r)) continue;
        if (cp.sceneSampleCount) {
            const std::uint32_t scene = pb.pathScene[r.path];
            const std::uint32_t k = atomicAdd(cp.sceneSampleCount + scene, 1U);
            if (k < cp.sampleCapacity) cp.samples[static_cast<std::size_t>(scene) * cp.sampleCapacity + k] = r;
        } else {
            const std::uint32_t k = atomicAdd(cp.sampleCount, 1U);
            if (k < cp.sampleCapacity) cp.samples[k] = r;
        }
    }
}

// The same gather in record order (deterministic sample order, so that a
// mini-batch holds the same samples run after run): valid records are
// counted per block of kGatherBlock, the blocks' bases follow the samples
// already gathered, then each block writes its samples in record order.
constexpr unsigned kGatherBlock = 1024;

__global__ __launch_bounds__(kGatherBlock) void gatherCountKernel(PathBuffers pb, CacheParams cp, std::uint32_t records,
                                                                  std::uint32_t* blockCount)
{
    const std::uint32_t i = blockIdx.x * kGatherBlock + threadIdx.x;
    RcQuery r;
    const bool ok = i < records && trainingSample(pb, cp, i, r);
    const int n = __syncthreads_count(ok);
    if (threadIdx.x == 0) blockCount[blockIdx.x] = static_cast<std::uint32_t>(n);
}

__global__ void gatherBasesKernel(CacheParams cp, std::uint32_t blocks, std::uint32_t* blockCount)
{
    std::uint32_t at = *cp.sampleCount;
    for (std::uint32_t b = 0; b < blocks; ++b) {
        const std::uint32_t n = blockCount[b];
        blockCount[b] = at;
        at += n;
    }
    *cp.sampleCount = at;
}

__global__ __launch_bounds__(kGatherBlock) void gatherWriteKernel(PathBuffers pb, CacheParams cp, std::uint32_t records,
                                                                  const std::uint32_t* blockBase)
{
    __shared__ std::uint32_t waveCount[kGatherBlock / 64U], waveBase[kGatherBlock / 64U];
    const std::uint32_t i = blockIdx.x * kGatherBlock + threadIdx.x;
    RcQuery r;
    const bool ok = i < records && trainingSample(pb, cp, i, r);
    const unsigned long long mask = __ballot(ok);
    const std::uint32_t lane = laneId(), wave = threadIdx.x >> 6;
    if (lane == 0U) waveCount[wave] = static_cast<std::uint32_t>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0U) {
        std::uint32_t at = blockBase[blockIdx.x];
        for (unsigned k = 0; k < kGatherBlock / 64U; ++k) {
            waveBase[k] = at;
            at += waveCount[k];
        }
    }
    __syncthreads();
    if (ok) {
        const std::uint32_t k = waveBase[wave] + static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
        if (k < cp.sampleCapacity) cp.samples[k] = r;
    }
}

// After the samples are gathered: the records are consumed (never reused
// by a later chunk), and with `outputs` (the cache's answers to the
// records' queries) a training lane path whose record at the terminate
// vertex exists takes, in the image, the value a path that stopped there
// gets: the radiance up to that vertex plus its throughput times the
// cache's answer (the plain rest of the path trained the cache only).
__global__ void finishTrainingKernel(PathBuffers pb, CacheParams cp, std::uint32_t records, const float* outputs,
                                     const float* calibration)
{
    if (cp.recordCount) records = min(*cp.recordCount, cp.recordSlots) * cp.recordVertices;
    const float scale = calibration ? *calibration : 1.0f;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < records; i += gridDim.x * blockDim.x) {
        const RcQuery r = cp.records[i];
        if (r.lane > 2U) continue;
        cp.records[i].lane = 0xffffffffU;
        if (outputs && i % cp.recordVertices == cp.terminateBounce) {
            const float y = outputs[3U * i + r.lane];
            pb.radiance[r.path] = r.target + cp.recordThroughput[i] * (scale * fmaxf(y, 0.0f));
        }
    }
}

// After the queries of a bounce are answered: counts them (queryCount + 2,
// 64-bit) and empties the queue for the next bounce.
// Marks `records` training records empty (lane ~0): records that no
// training path of the chunk writes are then never gathered.
__global__ void clearRecordsKernel(RcQuery* records, std::uint32_t n)
{
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        records[i].lane = 0xffffffffU;
}

__global__ void consumeQueriesKernel(CacheParams cp)
{
    *reinterpret_cast<unsigned long long*>(cp.queryCount + 2) += min(*cp.queryCount, cp.queryCapacity);
    *cp.queryCount = 0U;
}

// ---------------------------------------------------------------------------
// film: per pixel of the chunk, sum over its samples of
// (sum over 3 lanes of sensor(lambda) x radiance) / 3 in FP64 (:1882-1884).
// ---------------------------------------------------------------------------
__global__ void filmKernel(PathBuffers pb, ChunkDesc c, double4* batch)
{
    const std::uint32_t pixels = c.p1 - c.p0;
    const std::uint32_t samples = c.s1 - c.s0;
    for (std::uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < pixels; k += gridDim.x * blockDim.x) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (std::uint32_t so = 0; so < samples; ++so) {
            const std::uint32_t base = (so * pixels + k) * 3U;
            double sample[3] = {0.0, 0.0, 0.0};
            for (std::uint32_t lane = 0; lane < 3; ++lane) {
                double sens[3];
                spectralSensor64(static_cast<double>(pb.wavelength[base + lane]), sens);
                const double L = static_cast<double>(pb.radiance[base + lane]);
                sample[0] += sens[0] * L;
                sample[1] += sens[1] * L;
                sample[2] += sens[2] * L;
            }
            acc[0] += sample[0] / 3.0;
            acc[1] += sample[1] / 3.0;
            acc[2] += sample[2] / 3.0;
        }
        double4 b = batch[c.p0 + k];
        b.x += acc[0];
        b.y += acc[1];
        b.z += acc[2];
        batch[c.p0 + k] = b;
    }
}

// Batch film: per tile pixel of the current wave, the same FP64 sum as
// filmKernel, into its view's batch buffer (each pixel is in one tile).
__global__ void filmBatchKernel(PathBuffers pb, BatchWaves bw)
{
    const std::uint32_t w = *bw.cursor;
    const std::uint32_t ta = bw.waveTiles[w], tb = bw.waveTiles[w + 1];
    const std::uint32_t pixels = bw.wavePixels[w];
    for (std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < pixels; q += gridDim.x * blockDim.x) {
        const BatchTile& t = bw.tiles[findTile(bw.tiles, ta, tb, q, [](const BatchTile& x) { return x.pixelOffset; })];
        const BatchView& v = bw.views[t.scene];
        const std::uint32_t k = q - t.pixelOffset, tilePixels = t.p1 - t.p0;
        double acc[3] = {0.0, 0.0, 0.0};
        for (std::uint32_t so = 0; so < v.samples; ++so) {
            const std::uint32_t base = t.pathOffset + (so * tilePixels + k) * 3U;
            double sample[3] = {0.0, 0.0, 0.0};
            for (std::uint32_t lane = 0; lane < 3; ++lane) {
                double sens[3];
                spectralSensor64(static_cast<double>(pb.wavelength[base + lane]), sens);
                const double L = static_cast<double>(pb.radiance[base + lane]);
                sample[0] += sens[0] * L;
                sample[1] += sens[1] * L;
                sample[2] += sens[2] * L;
            }
            acc[0] += sample[0] / 3.0;
            acc[1] += sample[1] / 3.0;
            acc[2] += sample[2] / 3.0;
        }
        double4 b = v.batch[t.p0 + k];
        b.x += acc[0];
        b.y += acc[1];
        b.z += acc[2];
        v.batch[t.p0 + k] = b;
    }
}

// Moves the wave cursor on (last node of a wave's graph).
__global__ void advanceWaveKernel(BatchWaves bw, std::uint32_t* clear)
{
    *bw.cursor += 1U;
    if (clear) *clear = 0U;
}

// Progressive mean (:1886-1890) in FP64; alpha = 1.
__global__ void resolveKernel(double4* mean, const double4* batch, float4* out, std::uint32_t pixels,
                              std::uint32_t sampleIndex, std::uint32_t sampleCount)
{
    const double prevW = static_cast<double>(sampleIndex);
    const double total = prevW + static_cast<double>(sampleCount);
    for (std::uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += gridDim.x * blockDim.x) {
        const double4 prev = mean[p];
        const double4 b = batch[p];
        double4 m;
        m.x = (prev.x * prevW + b.x) / total;
        m.y = (prev.y * prevW + b.y) / total;
        m.z = (prev.z * prevW + b.z) / total;
        m.w = 1.0;
        mean[p] = m;
        out[p] = make_float4(static_cast<float>(m.x), static_cast<float>(m.y), static_cast<float>(m.z), 1.0f);
    }
}

// UpscaleNearest (src/hydra/render_pass.cpp:45-68) of an RGBA32F film: one
// thread per destination pixel, the reference's unsigned 32-bit index
// arithmetic.
__global__ void upscaleNearestKernel(const float4* src, std::uint32_t sw, std::uint32_t sh, float4* dst, std::uint32_t w,
                                     std::uint32_t h)
{
    const std::size_t n = static_cast<std::size_t>(w) * h;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        const std::uint32_t y = static_cast<std::uint32_t>(i / w), x = static_cast<std::uint32_t>(i % w);
        const std::uint32_t sy = min(sh - 1U, y * sh / h), sx = min(sw - 1U, x * sw / w);
        dst[i] = src[static_cast<std::size_t>(sy) * sw + sx];
    }
}

__global__ void zeroDoubleKernel(double4* p, std::uint32_t n)
{
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// Ray-query kernels behind hdcx_trace_closest / hdcx_trace_occluded.
template <bool k2L>
__global__ __launch_bounds__(kBlock) void queryClosestKernel(DeviceScene s, const float4* rays, std::uint32_t n, float4* hits)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        const RayPre r = prepareRay(a.x, a.y, a.z, b.x, b.y, b.z);
        HitRec h;
        const bool found = traceClosestS<false, k2L>(s, r, a.w, b.w, 0.0f, 0xffffffffU, h, st, nullptr);
        hits[i] = found ? make_float4(h.t, __uint_as_float(h.prim), h.u, h.v)
                        : make_float4(0.0f, __uint_as_float(0xffffffffU), 0.0f, 0.0f);
    }
}

template <bool k2L>
__global__ __launch_bounds__(kBlock) void queryAnyKernel(DeviceScene s, const float4* rays, std::uint32_t n, unsigned char* out)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        const RayPre r = prepareRay(a.x, a.y, a.z, b.x, b.y, b.z);
        out[i] = traceAnyS<false, k2L>(s, r, a.w, b.w, st, nullptr) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
hipError_t LaunchRaygen(const PathBuffers& pb, const ChunkDesc& c, std::uint32_t paths, const CacheParams& cp,
                        unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(raygenKernel, dim3(GridFor(paths, maxBlocks)), dim3(kBlock), 0, st, pb, c, paths, cp);
    return hipGetLastError();
}
std::size_t GatherScratchWords(std::uint32_t records) { return (records + kGatherBlock - 1U) / kGatherBlock + 1U; }

hipError_t LaunchGatherTraining(const PathBuffers& pb, const CacheParams& cp, std::uint32_t records, unsigned maxBlocks,
                                hipStream_t st, const float* outputs, const float* calibration, std::uint32_t* scratch)
{
    if (scratch && !cp.sceneSampleCount && !cp.recordCount) {
        const std::uint32_t blocks = (records + kGatherBlock - 1U) / kGatherBlock;
        if (blocks > 0) {
            hipLaunchKernelGGL(gatherCountKernel, dim3(blocks), dim3(kGatherBlock), 0, st, pb, cp, records, scratch);
            hipLaunchKernelGGL(gatherBasesKernel, dim3(1), dim3(1), 0, st, cp, blocks, scratch);
            hipLaunchKernelGGL(gatherWriteKernel, dim3(blocks), dim3(kGatherBlock), 0, st, pb, cp, records, scratch);
        }
    } else {
        hipLaunchKernelGGL(gatherTrainingKernel, dim3(GridFor(records, maxBlocks)), dim3(kBlock), 0, st, pb, cp, records);
    }
    hipLaunchKernelGGL(finishTrainingKernel, dim3(GridFor(records, maxBlocks)), dim3(kBlock), 0, st, pb, cp, records, outputs,
                       calibration);
    return hipGetLastError();
}
template <class Scenes, bool k2L, bool kOpaque>
void LaunchExtendK(const Scenes& s, const PathBuffers& pb, std::uint32_t in, int sharing, std::uint32_t maxWork,
                   unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    if (ctr)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(extendKernel<true, Scenes, k2L, kOpaque>),
                           dim3(StageGrid(extendKernel<true, Scenes, k2L, kOpaque>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, in, sharing, ctr);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(extendKernel<false, Scenes, k2L, kOpaque>),
                           dim3(StageGrid(extendKernel<false, Scenes, k2L, kOpaque>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, in, sharing, ctr);
}
template <class Scenes, bool k2L>
hipError_t LaunchExtendT(const Scenes& s, bool opaque, const PathBuffers& pb, std::uint32_t in, int sharing,
                         std::uint32_t maxWork, unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    if constexpr (k2L) {
        if (opaque) {
            LaunchExtendK<Scenes, k2L, true>(s, pb, in, sharing, maxWork, maxBlocks, ctr, st);
            return hipGetLastError();
        }
    } else {
        if (opaque) {
            if (ctr)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(extendRefillKernel<true, Scenes>),
                                   dim3(StageGrid(extendRefillKernel<true, Scenes>, maxWork, maxBlocks)), dim3(kBlock), 0,
                                   st, s, pb, in, sharing, ctr);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(extendRefillKernel<false, Scenes>),
                                   dim3(StageGrid(extendRefillKernel<false, Scenes>, maxWork, maxBlocks)), dim3(kBlock), 0,
                                   st, s, pb, in, sharing, ctr);
            return hipGetLastError();
        }
    }
    LaunchExtendK<Scenes, k2L, false>(s, pb, in, sharing, maxWork, maxBlocks, ctr, st);
    return hipGetLastError();
}
template <class Scenes, bool k2L, bool kOpaque>
void LaunchShadowK(const Scenes& s, const PathBuffers& pb, bool shared, std::uint32_t maxWork, unsigned maxBlocks,
                   TraversalCounters* ctr, hipStream_t st)
{
    if (ctr)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowKernel<true, Scenes, k2L, kOpaque>),
                           dim3(StageGrid(shadowKernel<true, Scenes, k2L, kOpaque>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, shared, ctr);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowKernel<false, Scenes, k2L, kOpaque>),
                           dim3(StageGrid(shadowKernel<false, Scenes, k2L, kOpaque>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, shared, ctr);
}
template <class Scenes, bool k2L>
hipError_t LaunchShadowT(const Scenes& s, bool opaque, const PathBuffers& pb, bool shared, std::uint32_t maxWork,
                         unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    if (opaque) LaunchShadowK<Scenes, k2L, true>(s, pb, shared, maxWork, maxBlocks, ctr, st);
    else LaunchShadowK<Scenes, k2L, false>(s, pb, shared, maxWork, maxBlocks, ctr, st);
    return hipGetLastError();
}
hipError_t LaunchExtend(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int sharing, std::uint32_t maxWork,
                        unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    const bool opaque = s.opaque != 0U;
    return s.instanceCount
               ? LaunchExtendT<SingleScene, true>(SingleScene{s}, opaque, pb, in, sharing, maxWork, maxBlocks, ctr, st)
               : LaunchExtendT<SingleScene, false>(SingleScene{s}, opaque, pb, in, sharing, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchShadow(const DeviceScene& s, const PathBuffers& pb, bool shared, std::uint32_t maxWork, unsigned maxBlocks,
                        TraversalCounters* ctr, hipStream_t st)
{
    const bool opaque = s.opaque != 0U;
    return s.instanceCount
               ? LaunchShadowT<SingleScene, true>(SingleScene{s}, opaque, pb, shared, maxWork, maxBlocks, ctr, st)
               : LaunchShadowT<SingleScene, false>(SingleScene{s}, opaque, pb, shared, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchExtendBatch(const DeviceScene* table, bool twoLevel, bool opaque, const PathBuffers& pb,
                             std::uint32_t in, int sharing, std::uint32_t maxWork, unsigned maxBlocks,
                             TraversalCounters* ctr, hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchExtendT<SceneTable, true>(t, opaque, pb, in, sharing, maxWork, maxBlocks, ctr, st)
                    : LaunchExtendT<SceneTable, false>(t, opaque, pb, in, sharing, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchShadowBatch(const DeviceScene* table, bool twoLevel, bool opaque, const PathBuffers& pb,
                             std::uint32_t maxWork, unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchShadowT<SceneTable, true>(t, opaque, pb, true, maxWork, maxBlocks, ctr, st)
                    : LaunchShadowT<SceneTable, false>(t, opaque, pb, true, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchRaygenBatch(const PathBuffers& pb, const BatchWaves& bw, const CacheParams& cp, std::uint32_t maxWork,
                             unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(raygenBatchKernel, dim3(GridFor(maxWork, maxBlocks)), dim3(kBlock), 0, st, pb, bw, cp);
    return hipGetLastError();
}
hipError_t LaunchFilmBatch(const PathBuffers& pb, const BatchWaves& bw, std::uint32_t maxWork, unsigned maxBlocks,
                           hipStream_t st)
{
    hipLaunchKernelGGL(filmBatchKernel, dim3(GridFor(maxWork, maxBlocks)), dim3(kBlock), 0, st, pb, bw);
    return hipGetLastError();
}
hipError_t LaunchAdvanceWave(const BatchWaves& bw, std::uint32_t* clear, hipStream_t st)
{
    hipLaunchKernelGGL(advanceWaveKernel, dim3(1), dim3(1), 0, st, bw, clear);
    return hipGetLastError();
}
hipError_t LaunchFilm(const PathBuffers& pb, const ChunkDesc& c, double4* batch, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(filmKernel, dim3(GridFor(c.p1 - c.p0, maxBlocks)), dim3(kBlock), 0, st, pb, c, batch);
    return hipGetLastError();
}
hipError_t LaunchClearRecords(RcQuery* records, std::uint32_t n, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(clearRecordsKernel, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, records, n);
    return hipGetLastError();
}
hipError_t LaunchConsumeQueries(const CacheParams& cp, hipStream_t st)
{
    hipLaunchKernelGGL(consumeQueriesKernel, dim3(1), dim3(1), 0, st, cp);
    return hipGetLastError();
}
hipError_t LaunchResolve(double4* mean, const double4* batch, float4* out, std::uint32_t pixels, std::uint32_t sampleIndex,
                         std::uint32_t sampleCount, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(resolveKernel, dim3(GridFor(pixels, maxBlocks)), dim3(kBlock), 0, st, mean, batch, out, pixels,
                       sampleIndex, sampleCount);
    return hipGetLastError();
}
hipError_t LaunchUpscaleNearest(const float4* src, std::uint32_t sw, std::uint32_t sh, float4* dst, std::uint32_t w,
                                std::uint32_t h, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(upscaleNearestKernel, dim3(GridFor(w * h, maxBlocks)), dim3(kBlock), 0, st, src, sw, sh, dst, w, h);
    return hipGetLastError();
}
hipError_t LaunchZero(double4* p, std::uint32_t n, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(zeroDoubleKernel, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, p, n);
    return hipGetLastError();
}
hipError_t LaunchQueryClosest(const DeviceScene& s, const float4* rays, std::uint32_t n, float4* hits, unsigned maxBlocks,
                              hipStream_t st)
{
    if (s.instanceCount)
        hipLaunchKernelGGL(queryClosestKernel<true>, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, s, rays, n, hits);
    else
        hipLaunchKernelGGL(queryClosestKernel<false>, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, s, rays, n, hits);
    return hipGetLastError();
}
hipError_t LaunchQueryAny(const DeviceScene& s, const float4* rays, std::uint32_t n, unsigned char* out, unsigned maxBlocks,
                          hipStream_t st)
{
    if (s.instanceCount)
        hipLaunchKernelGGL(queryAnyKernel<true>, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, s, rays, n, out);
    else
        hipLaunchKernelGGL(queryAnyKernel<false>, dim3(GridFor(n, maxBlocks)), dim3(kBlock), 0, st, s, rays, n, out);
    return hipGetLastError();
}

} // namespace hdcx
