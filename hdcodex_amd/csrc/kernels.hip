// Wavefront spectral path tracer for MI355X (gfx950).
//
// One chunk of lane paths (3 per pixel sample) flows through
//   raygen -> [extend -> shade -> shadow] x maxBounces -> film
// with every stage a grid-stride kernel over a device-side queue whose
// length the previous stage wrote; the host never reads a count back.
// shade compacts surviving paths and paths with pending shadow rays with a
// wave ballot + prefix popcount and one atomic per wave.
//
// FP contraction is off for the whole integrator: every instantiation of a
// stage (one scene or a scene table, one or two levels, walks inline or
// deferred) then rounds the same operations the same way, and the FP32
// arithmetic follows the oracle's (built with -ffp-contract=off).
#pragma clang fp contract(off)
#include "kernel_api.h"
#include "dev_shading.h"
#include "dev_traverse2.h"
#include "rcache.h"

namespace hdcx {

constexpr int kBlock = 128;

__device__ __forceinline__ std::uint32_t laneId() { return __lane_id(); }

// Appends `item` for lanes with `keep` to `queue` with one atomic per wave.
__device__ __forceinline__ void waveAppend(bool keep, std::uint32_t item, std::uint32_t* queue, std::uint32_t* counter)
{
    const unsigned long long mask = __ballot(keep);
    if (mask == 0ULL) return;
    const unsigned long long active = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(active)) - 1;
    const std::uint32_t lane = laneId();
    const std::uint32_t total = static_cast<std::uint32_t>(__popcll(mask));
    const std::uint32_t before = static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
    std::uint32_t base = 0;
    if (static_cast<int>(lane) == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    if (keep) queue[base + before] = item;
}

// Index of this lane's entry in a compacted array (one atomic per wave), or
// 0xffffffff when it does not want one.
__device__ __forceinline__ std::uint32_t waveSlot(bool want, std::uint32_t* counter)
{
    const unsigned long long mask = __ballot(want);
    if (mask == 0ULL) return 0xffffffffU;
    const unsigned long long active = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(active)) - 1;
    const std::uint32_t lane = laneId();
    std::uint32_t base = 0;
    if (static_cast<int>(lane) == leader) base = atomicAdd(counter, static_cast<std::uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    return want ? base + static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL))) : 0xffffffffU;
}

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
    // A single render clears the training slot this path will write. (In a
    // batch, slots are handed out per wave by raygenBatchKernel and the
    // gather of the previous wave left every handed-out record empty.)
    if (cp.enabled && !cp.trainSlot && (ps + cp.trainOffset) % cp.trainStride == 0U) {
        const std::uint32_t slot = (ps + cp.trainOffset) / cp.trainStride;
        if (slot < cp.recordSlots)
            for (std::uint32_t j = 0; j < cp.recordVertices; ++j)
                cp.records[(slot * 3U + lane) * cp.recordVertices + j].lane = 0xffffffffU;
    }
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
    pb.flags[pid] = kEnvOnMiss | kPrevDelta;
    pb.mis[pid] = make_float2(0.0f, 0.0f);
    pb.medium[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int k = 0; k < kShadowSlots; ++k) pb.shadowLight[k * pb.capacity + pid] = 0xffffffffU;
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

template <bool k2L>
__device__ __forceinline__ float materialOpacity(const DeviceScene& s, std::uint32_t prim, float bu, float bv)
{
    const TriRef tr = resolvePrim<k2L>(s, prim);
    const GpuMaterialRecord& m = s.materials[triangleMaterial(s, tr)];
    const std::uint32_t tri = tr.tri;
    const float2 t0 = s.texcoords[tri * 3], t1 = s.texcoords[tri * 3 + 1], t2 = s.texcoords[tri * 3 + 2];
    const float w0 = 1.0f - bu - bv;
    const float u = t0.x * w0 + t1.x * bu + t2.x * bv, v = t0.y * w0 + t1.y * bu + t2.y * bv;
    return clampf(matScalar(s, m.textureIndices1.x, u, v, m.transmissionOpacityIor.y), 0.0f, 1.0f);
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
// once, for lane 0. kShareLanes: pairKernel queued the paths that trace
// (pb.extendQueue) with a mask of the later lanes of their sample whose
// next ray it found equal (pb.follow). The hit (and, in a scene with
// cut-outs, the random state after the opacity draws) is written to the
// path and to the lanes that share its ray: the same values each lane
// would compute for itself.
// ---------------------------------------------------------------------------
template <bool kCount, class Scenes, bool k2L>
__global__ __launch_bounds__(kBlock) void extendKernel(Scenes scenes, PathBuffers pb, std::uint32_t in, int sharing,
                                                      TraversalCounters* ctr)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pb.counters[(in ^ 1U) * kCounterStride] = 0U;
        pb.counters[2 * kCounterStride] = 0U;
        pb.counters[3 * kCounterStride] = 0U;
        pb.counters[(4U + (in ^ 1U)) * kCounterStride] = 0U;
        pb.counters[6 * kCounterStride] = 0U;
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
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rays; i += gridDim.x * blockDim.x) {
        const std::uint32_t pid = camera ? 3U * i : queue[i];
        const std::uint32_t followers = camera ? 3U : leaders ? pb.follow[pid] : 0U;
        const DeviceScene& s = scenes.Get(pid);
        const float4 o = pb.rayO[pid], d = pb.rayD[pid];
        const RayPre r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
        HitRec h;
        bool found;
        std::uint32_t rng = 0U;
        if (s.opaque) {
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
        if (!s.opaque) pb.rng[pid] = rng;
        for (std::uint32_t k = 0; k < 2U; ++k) {
            if (!(followers & (1U << k))) continue;
            pb.hit[pid + 1U + k] = hit;
            if (!s.opaque) pb.rng[pid + 1U + k] = rng;
        }
    }
}

// ---------------------------------------------------------------------------
// pair: lane sharing for the next extend (kShareLanes), over the paths that
// survived this bounce (queue[out]) in queue order. With the same random
// numbers, the lanes of a pixel sample keep the same rays until a lobe or
// light choice differs. A path whose next ray (and, in a scene with
// cut-outs, random state) equals bit for bit that of an earlier lane of its
// sample, held by one of the two preceding threads of the wave, is not
// traced: that lane's extend writes the hit to both. Lanes of a sample have
// consecutive path ids (lane = pid % 3) and keep their order in the queues,
// but another path can sit between two of them (the queues are filled wave
// by wave in atomic order), so the earliest equal lane may be out of a
// path's sight while the lane it sees follows that earlier one: the
// earliest visible equal lane is taken, and kept only if it traces itself,
// so every path that does not trace has a leader that does. The tracing
// paths go to extendQueue[out], one atomic per block of 16 waves, with the
// mask of their followers in follow[].
// ---------------------------------------------------------------------------
constexpr unsigned kPairBlock = 1024;

// Leader / follower decision of one path among the lanes of its pixel
// sample, from kWords words that must be equal bit for bit (see pairKernel).
// Returns the follower mask of a path that traces (bit j: pid + 1 + j) in
// `followers`; true when the path traces.
template <int kWords>
__device__ __forceinline__ bool pairLanes(bool live, std::uint32_t pid, const std::uint32_t (&w)[kWords],
                                          std::uint32_t& followers)
{
    const std::uint32_t lane = laneId(), sampleLane = pid % 3U;
    std::uint32_t back = 0U;  // 0: traced here; 1 or 2: path pid - back traces it
    std::uint32_t dist = 0U;  // threads back to that path
    for (std::uint32_t k = 1U; k <= 2U; ++k) {
        const std::uint32_t qp = __shfl_up(pid, k);
        const int ql = __shfl_up(static_cast<int>(live), k);
        bool same = live && lane >= k && ql != 0 && qp < pid && pid - qp <= sampleLane;
        for (int j = 0; j < kWords; ++j) same = (__shfl_up(w[j], k) == w[j]) && same;
        if (same && pid - qp > back) {
            back = pid - qp;
            dist = k;
        }
    }
    const std::uint32_t leaderBack1 = __shfl_up(back, 1U), leaderBack2 = __shfl_up(back, 2U);
    if (back != 0U && (dist == 1U ? leaderBack1 : leaderBack2) != 0U) back = 0U;
    followers = 0U;
    for (std::uint32_t k = 1U; k <= 2U; ++k) {
        const std::uint32_t qp = __shfl_down(pid, k);
        const std::uint32_t qb = __shfl_down(back, k);
        if (lane + k < 64U && qb != 0U && qp - qb == pid) followers |= 1U << (qp - pid - 1U);
    }
    return live && back == 0U;
}

// Appends pid for threads with `want` to `queue`: one atomic per block of
// kPairBlock threads. Every thread of the block must call it.
__device__ __forceinline__ void pairAppend(bool want, std::uint32_t pid, std::uint32_t* queue, std::uint32_t* counter,
                                           std::uint32_t* waveCount, std::uint32_t* waveBase)
{
    constexpr unsigned kWaves = kPairBlock / 64U;
    const std::uint32_t lane = laneId(), wave = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(want);
    if (lane == 0U) waveCount[wave] = static_cast<std::uint32_t>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0U) {
        std::uint32_t total = 0U;
        for (unsigned k = 0; k < kWaves; ++k) total += waveCount[k];
        std::uint32_t at = total ? atomicAdd(counter, total) : 0U;
        for (unsigned k = 0; k < kWaves; ++k) {
            waveBase[k] = at;
            at += waveCount[k];
        }
    }
    __syncthreads();
    if (want) queue[waveBase[wave] + static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)))] = pid;
    __syncthreads();
}

template <class Scenes>
__global__ __launch_bounds__(kPairBlock) void pairKernel(Scenes scenes, PathBuffers pb, std::uint32_t out)
{
    __shared__ std::uint32_t waveCount[kPairBlock / 64U], waveBase[kPairBlock / 64U];
    const std::uint32_t count = pb.counters[out * kCounterStride];
    const std::uint32_t* queue = pb.queue[out];
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (count + stride - 1U) / stride * stride;  // every thread reaches the barriers
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < padded; i += stride) {
        const bool live = i < count;
        std::uint32_t pid = 0U;
        std::uint32_t w[7] = {0U, 0U, 0U, 0U, 0U, 0U, 0U};  // ray origin, direction, random state (cut-outs only)
        if (live) {
            pid = queue[i];
            const float4 o = pb.rayO[pid], d = pb.rayD[pid];
            w[0] = __float_as_uint(o.x);
            w[1] = __float_as_uint(o.y);
            w[2] = __float_as_uint(o.z);
            w[3] = __float_as_uint(d.x);
            w[4] = __float_as_uint(d.y);
            w[5] = __float_as_uint(d.z);
            if (!scenes.Get(pid).opaque) w[6] = pb.rng[pid];
        }
        std::uint32_t followers;
        const bool traces = pairLanes<7>(live, pid, w, followers);
        if (traces) pb.follow[pid] = followers;
        pairAppend(traces, pid, pb.extendQueue[out], pb.counters + (4U + out) * kCounterStride, waveCount, waveBase);
    }
}

// shadowPair: the same grouping over the paths with shadow rays
// (shadowQueue), comparing every slot's light, origin, extent and
// direction. The shadow kernel then traces each group's rays once and
// applies the visibility to every lane of the group with the lane's own
// wavelength and contribution (the visibility test itself draws no random
// numbers). Leaders go to shadowLeaders with their masks in shadowFollow.
template <class Scenes>
__global__ __launch_bounds__(kPairBlock) void shadowPairKernel(Scenes scenes, PathBuffers pb)
{
    __shared__ std::uint32_t waveCount[kPairBlock / 64U], waveBase[kPairBlock / 64U];
    const std::uint32_t count = pb.counters[2 * kCounterStride];
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (count + stride - 1U) / stride * stride;  // every thread reaches the barriers
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < padded; i += stride) {
        const bool live = i < count;
        std::uint32_t pid = 0U;
        std::uint32_t w[8 * kShadowSlots];
        for (int j = 0; j < 8 * kShadowSlots; ++j) w[j] = 0U;
        if (live) {
            pid = pb.shadowQueue[i];
            for (int k = 0; k < kShadowSlots; ++k) {
                const std::size_t idx = static_cast<std::size_t>(k) * pb.capacity + pid;
                const std::uint32_t light = pb.shadowLight[idx];
                w[8 * k] = light;
                if (light == 0xffffffffU) break;
                const float4 o = pb.shadowO[idx], d = pb.shadowD[idx];
                w[8 * k + 1] = __float_as_uint(o.x);
                w[8 * k + 2] = __float_as_uint(o.y);
                w[8 * k + 3] = __float_as_uint(o.z);
                w[8 * k + 4] = __float_as_uint(o.w);
                w[8 * k + 5] = __float_as_uint(d.x);
                w[8 * k + 6] = __float_as_uint(d.y);
                w[8 * k + 7] = __float_as_uint(d.z);
            }
        }
        std::uint32_t followers;
        const bool traces = pairLanes<8 * kShadowSlots>(live, pid, w, followers);
        if (traces) pb.shadowFollow[pid] = followers;
        pairAppend(traces, pid, pb.shadowLeaders, pb.counters + 6U * kCounterStride, waveCount, waveBase);
    }
}

// Closest opaque hit used by the subsurface walk (:1130-1151).
template <bool k2L>
__device__ bool opaqueBoundary(const DeviceScene& s, V3 o, V3 d, TraversalStack& st, V3& boundary, V3& bn, float& dist)
{
    const RayPre r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
    HitRec h;
    if (!traceClosestS<false, k2L>(s, r, 0.001f, 10000.0f, 0.0f, 0xffffffffU, h, st, nullptr)) return false;
    dist = h.t;
    boundary = o + d * dist;
    V3 p0, p1, p2;
    triangleWorld(s, resolvePrim<k2L>(s, h.prim), p0, p1, p2);
    bn = normalize(cross(p1 - p0, p2 - p0));
    if (dot(bn, d) > 0.0f) bn = -bn;
    return true;
}

// randomWalkSubsurface (:1153-1186)
template <bool k2L>
__device__ bool randomWalk(const DeviceScene& s, V3 entry, V3 outN, float mfp, float albedo, float aniso,
                           std::uint32_t& rng, TraversalStack& st, V3& exitO, V3& exitD, float& atten)
{
    const float radius = fmaxf(mfp, 0.001f);
    const float sigmaS = 1.0f / radius;
    const float sigmaA = -logf(clampf(albedo, 0.001f, 0.9999f)) / radius;
    V3 o = entry - outN * 0.003f;
    V3 d = cosineHemisphere(-outN, rng);
    atten = 1.0f;
    for (int step = 0; step < 8; ++step) {
        V3 boundary, bn;
        float bd;
        if (!opaqueBoundary<k2L>(s, o, d, st, boundary, bn, bd)) return false;
        const float scatter = -logf(fmaxf(1.0f - randomFloat(rng), 1e-6f)) / sigmaS;
        const bool force = step == 7;
        const float travel = force ? bd : fminf(scatter, bd);
        atten *= expf(-sigmaA * travel);
        if (force || scatter >= bd) {
            exitD = d;
            exitO = boundary + d * 0.003f;
            return true;
        }
        o = o + d * scatter;
        d = sampleHG(d, aniso, rng);
    }
    return false;
}

__device__ __forceinline__ void queueShadow(const PathBuffers& pb, std::uint32_t pid, int& slot, V3 o, V3 d,
                                            float tmax, float contribution, std::uint32_t light)
{
    const std::size_t k = static_cast<std::size_t>(slot) * pb.capacity + pid;
    pb.shadowO[k] = make_float4(o.x, o.y, o.z, tmax);
    pb.shadowD[k] = make_float4(d.x, d.y, d.z, contribution);
    pb.shadowLight[k] = light;
    ++slot;
}

// ---------------------------------------------------------------------------
// shade: one path vertex (:1245-1879). Radiance that needs a visibility
// test is queued as a shadow ray carrying its unoccluded contribution; the
// shadow stage adds it in slot order, which keeps the reference's order of
// additions (emission, analytic NEE, mesh-light NEE).
//
// Subsurface walks (kWalk): kShadeWalkInline runs each walk here.
// kShadeWalkDefer stops a path at the walk, stores the walk's inputs in a
// walk-queue slot and commits none of the path's state (what it did write,
// the shadow-ray slots and training records, is rewritten with the same
// values later); walkKernel then runs the walks without shade's register
// load, and kShadeWalkResume shades the queued paths again from their
// unchanged state, taking the walk's outcome and random state from the
// slot. The three steps give the same results as kShadeWalkInline.
// ---------------------------------------------------------------------------
template <class Scenes, bool k2L, int kWalk>
__global__ __launch_bounds__(kBlock) void shadeKernel(Scenes scenes, PathBuffers pb, std::uint32_t in, int bounce,
                                                     int lastBounce, CacheParams cp)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
    const std::uint32_t count =
        kWalk == kShadeWalkResume ? pb.counters[3 * kCounterStride] : pb.counters[in * kCounterStride];
    const std::uint32_t* queue = kWalk == kShadeWalkResume ? pb.walkQueue : pb.queue[in];
    std::uint32_t* outQueue = pb.queue[in ^ 1U];
    std::uint32_t* outCounter = pb.counters + (in ^ 1U) * kCounterStride;
    // Every lane of a wave must reach the ballots: iterate over the padded range.
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (count + stride - 1U) / stride * stride;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < padded; i += stride) {
        bool keep = false;
        int slot = 0;
        std::uint32_t pid = 0;
        bool wantQuery = false;
        RcQuery query{};
        bool deferred = false;
        float4 walkIn0{}, walkIn1{}, walkIn2{};
        if (i < count) {
            pid = queue[i];
            const DeviceScene& s = scenes.Get(pid);
            const float4 hit4 = pb.hit[pid];
            const std::uint32_t prim = __float_as_uint(hit4.y);
            float4 ro4 = pb.rayO[pid], rd4 = pb.rayD[pid];
            V3 ro = v3(ro4), rd = v3(rd4);
            std::uint32_t rng = pb.rng[pid];
            const float wl = pb.wavelength[pid];
            float throughput = pb.throughput[pid];
            float radiance = pb.radiance[pid];
            std::uint32_t flags = pb.flags[pid];
            float2 mis = pb.mis[pid];
            float4 med = pb.medium[pid];
            float2 spreadOut = make_float2(0.0f, 0.0f);
            bool spreadSet = false;
            for (int k = 0; k < kShadowSlots; ++k) pb.shadowLight[k * pb.capacity + pid] = 0xffffffffU;

            if (prim == 0xffffffffU) {
                if (flags & kEnvOnMiss) {
                    const float w = (flags & kPrevDelta) ? 1.0f : powerHeuristic(mis.x, mis.y);
                    radiance += throughput * w * illuminantSpectrum(environmentRgb(s, rd), wl);
                }
                pb.radiance[pid] = radiance;
            } else {
                const float distance = hit4.x, bu = hit4.z, bv = hit4.w;
                bool mediumScattered = false;
                if (flags & kInMedium) {
                    if (med.y > 0.0f) {
                        const float sd = -logf(fmaxf(1.0f - randomFloat(rng), 1e-6f)) / med.y;
                        if (sd < distance) {
                            throughput *= expf(-med.x * sd);
                            ro = ro + rd * sd;
                            rd = sampleHG(rd, med.z, rng);
                            flags |= kEnvOnMiss | kPrevDelta;
                            mis.x = 0.0f;
                            mediumScattered = true;
                            keep = true;
                        }
                    }
                    if (!mediumScattered) throughput *= expf(-med.x * distance);
                }
                if (!mediumScattered) {
                    const TriRef tr = resolvePrim<k2L>(s, prim);
                    V3 p0, p1, p2;
                    triangleWorld(s, tr, p0, p1, p2);
                    const V3 gn = normalize(cross(p1 - p0, p2 - p0));
                    const bool front = dot(gn, rd) < 0.0f;
                    const V3 ogn = front ? gn : -gn;
                    const float w0 = 1.0f - bu - bv;
                    V3 normal;
                    {
                        const float4 n0 = cornerNormal(s, tr, 0), n1 = cornerNormal(s, tr, 1), n2 = cornerNormal(s, tr, 2);
                        V3 an = v3(n0) * w0 + v3(n1) * bu + v3(n2) * bv;
                        if (dot(an, an) < 1e-12f) an = gn;
                        else {
                            an = normalize(an);
                            if (dot(an, gn) < 0.0f) an = -an;
                        }
                        normal = front ? an : -an;
                    }
                    const float2 t0 = s.texcoords[tr.tri * 3], t1 = s.texcoords[tr.tri * 3 + 1], t2 = s.texcoords[tr.tri * 3 + 2];
                    const float su = t0.x * w0 + t1.x * bu + t2.x * bv, sv = t0.y * w0 + t1.y * bu + t2.y * bv;
                    const V3 hp = ro + rd * distance;
                    const GpuMaterialRecord& m = s.materials[triangleMaterial(s, tr)];

                    // Material fetch (:1296-1437).
                    V3 baseRgb = v3(m.baseColorMetalness);
                    if (m.textureIndices0.x != kMissingTexture) {
                        const float4 smp = sampleMaterialTexture(s, m.textureIndices0.x, su, sv,
                                                                 make_float4(baseRgb.x, baseRgb.y, baseRgb.z, 1.0f));
                        const int channel = static_cast<int>(rintf(m.baseColorTextureScale.w));
                        const float comps[4] = {smp.x, smp.y, smp.z, smp.w};
                        const V3 src = channel < 0 ? v3(smp.x, smp.y, smp.z) : v3(comps[min(channel, 3)]);
                        baseRgb = v3(m.baseColorTextureBias) + v3(m.baseColorTextureScale) * src;
                    }
                    Closure c;
                    const float baseColor = reflectanceSpectrum(baseRgb, wl);
                    c.metalness = clampf(matScalar(s, m.textureIndices0.y, su, sv, m.baseColorMetalness.w), 0.0f, 1.0f);
                    {
                        float rx = m.emissionRoughness.w, ry = m.subsurfaceColor.w;
                        if (m.textureIndices0.z != kMissingTexture) {
                            const float4 r = sampleMaterialTexture(s, m.textureIndices0.z, su, sv, make_float4(rx, ry, 0.0f, 1.0f));
                            rx = r.x;
                            ry = r.y;
                        }
                        c.rx = clampf(rx, 0.02f, 1.0f);
                        c.ry = clampf(ry, 0.02f, 1.0f);
                    }
                    c.diffuseRoughness = clampf(matScalar(s, m.textureIndices4.z, su, sv, m.diffuseProperties.x), 0.0f, 1.0f);
                    c.diffuseWeight = clampf(matScalar(s, m.textureIndices4.w, su, sv, m.diffuseProperties.z), 0.0f, 1.0f);
                    c.diffuseModel = static_cast<std::uint32_t>(m.diffuseProperties.y + 0.5f);
                    const float2 uv0 = t0, uv1 = t1, uv2 = t2;
                    // applyNormalMap (:807-829)
                    if (m.textureIndices1.y != kMissingTexture) {
                        const float du1 = uv1.x - uv0.x, dv1 = uv1.y - uv0.y, du2 = uv2.x - uv0.x, dv2 = uv2.y - uv0.y;
                        const float det = du1 * dv2 - dv1 * du2;
                        if (fabsf(det) >= 1e-8f) {
                            const V3 e1 = p1 - p0, e2 = p2 - p0;
                            V3 t = normalize((e1 * dv2 - e2 * dv1) / det);
                            t = normalize(t - normal * dot(normal, t));
                            V3 b = normalize(cross(normal, t));
                            if (det < 0.0f) b = -b;
                            const float4 mp = sampleMaterialTexture(s, m.textureIndices1.y, su, sv, make_float4(0.5f, 0.5f, 1.0f, 1.0f));
                            float mx = mp.x * 2.0f - 1.0f, my = mp.y * 2.0f - 1.0f;
                            const float mz = mp.z * 2.0f - 1.0f;
                            if (m.coatWeightRoughnessIor.w > 0.5f) my = -my;
                            mx *= m.subsurfaceWeightScale.w;
                            my *= m.subsurfaceWeightScale.w;
                            const V3 r = normalize(t * mx + b * my + normal * mz);
                            normal = dot(r, normal) < 0.0f ? -r : r;
                        }
                    }
                    // surfaceTangent (:831-841)
                    {
                        const float du1 = uv1.x - uv0.x, dv1 = uv1.y - uv0.y, du2 = uv2.x - uv0.x, dv2 = uv2.y - uv0.y;
                        const float det = du1 * dv2 - dv1 * du2;
                        c.tangent = fabsf(det) < 1e-8f ? orthogonalTangent(normal, v3(1.0f, 0.0f, 0.0f))
                                                       : orthogonalTangent(normal, ((p1 - p0) * dv2 - (p2 - p0) * dv1) / det);
                    }
                    c.normal = normal;
                    // Emission with MIS against the emissive-triangle pdf (:1328-1337).
                    V3 emissionRgb;
                    {
                        const float weight = fmaxf(m.transmissionOpacityIor.w, 0.0f);
                        if (m.textureIndices0.w == kMissingTexture) emissionRgb = v3(m.emissionRoughness) * weight;
                        else {
                            const float4 e = sampleMaterialTexture(s, m.textureIndices0.w, su, sv, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                            emissionRgb = v3(e.x, e.y, e.z) * weight;
                        }
                    }
                    float emissionMis = 1.0f;
                    if (!(flags & kPrevDelta) && (emissionRgb.x > 0.0f || emissionRgb.y > 0.0f || emissionRgb.z > 0.0f)) {
                        const float ec = fabsf(dot(gn, -rd));
                        // Instanced triangles are never sampled as lights (pdf 0).
                        const float epdf = (!k2L || prim < s.drawnTriangles) ? s.emissive[prim].y : 0.0f;
                        const float lp = ec <= 0.0f ? 0.0f : epdf * (distance * distance) / ec;
                        emissionMis = powerHeuristic(mis.x, lp);
                    }
                    if (emissionRgb.x != 0.0f || emissionRgb.y != 0.0f || emissionRgb.z != 0.0f)
                        radiance += throughput * illuminantSpectrum(emissionRgb, wl) * emissionMis;

                    c.transmission = clampf(m.transmissionOpacityIor.x, 0.0f, 1.0f);
                    V3 transRgb = v3(m.transmissionColorThinWalled);
                    if (c.transmission > 0.0f) transRgb = matRgb(s, m.textureIndices1.z, su, sv, transRgb);
                    const float transmissionColor = reflectanceSpectrum(transRgb, wl);
                    c.subsurface = 0.0f;
                    c.diffuseColor = baseColor;
                    V3 sssColorRgb = v3(m.subsurfaceColor);
                    V3 sssRadiusRgb = vmax(v3(m.subsurfaceRadius), 0.001f);
                    if (m.subsurfaceWeightScale.x > 0.0f || m.textureIndices2.x != kMissingTexture) {
                        c.subsurface = clampf(matScalar(s, m.textureIndices2.x, su, sv, m.subsurfaceWeightScale.x), 0.0f, 1.0f);
                        if (c.subsurface > 0.0f) {
                            sssColorRgb = matRgb(s, m.textureIndices2.y, su, sv, v3(m.subsurfaceColor));
                            sssRadiusRgb = vmax(matRgb(s, m.textureIndices2.z, su, sv, v3(m.subsurfaceRadius)), 0.001f);
                            c.diffuseColor = mixf(baseColor, reflectanceSpectrum(sssColorRgb, wl), c.subsurface);
                        }
                    }
                    const float ior = dispersedIor(fmaxf(m.transmissionOpacityIor.z, 1.0001f), m.transmissionMedium.z,
                                                   m.transmissionMedium.w, wl);
                    const float specularWeight = clampf(matScalar(s, m.textureIndices3.x, su, sv, m.specularWeightColor.x), 0.0f, 1.0f);
                    const float specularColor = reflectanceSpectrum(
                        matRgb(s, m.textureIndices3.y, su, sv, v3(m.specularWeightColor.y, m.specularWeightColor.z, m.specularWeightColor.w)), wl);
                    float dielectric = (ior - 1.0f) / (ior + 1.0f);
                    dielectric *= dielectric;
                    c.f0 = mixf(clampf(dielectric * specularWeight * specularColor, 0.0f, 1.0f), baseColor, c.metalness);
                    c.coat = 0.0f;
                    c.crx = m.coatWeightRoughnessIor.y;
                    c.cry = m.subsurfaceRadius.w;
                    c.coatF0 = 0.0f;
                    if (m.coatWeightRoughnessIor.x > 0.0f || m.textureIndices3.z != kMissingTexture) {
                        c.coat = clampf(matScalar(s, m.textureIndices3.z, su, sv, m.coatWeightRoughnessIor.x), 0.0f, 1.0f);
                        if (c.coat > 0.0f) {
                            const float coatColor = reflectanceSpectrum(matRgb(s, m.textureIndices3.w, su, sv, v3(m.coatColor)), wl);
                            float crx = c.crx, cry = c.cry;
                            if (m.textureIndices4.x != kMissingTexture) {
                                const float4 r = sampleMaterialTexture(s, m.textureIndices4.x, su, sv, make_float4(crx, cry, 0.0f, 1.0f));
                                crx = r.x;
                                cry = r.y;
                            }
                            c.crx = clampf(crx, 0.02f, 1.0f);
                            c.cry = clampf(cry, 0.02f, 1.0f);
                            const float cior = fmaxf(m.coatWeightRoughnessIor.z, 1.0001f);
                            float cr = (cior - 1.0f) / (cior + 1.0f);
                            cr *= cr;
                            c.coatF0 = clampf(cr * coatColor, 0.0f, 1.0f);
                        }
                    }
                    c.sheen = clampf(matScalar(s, m.textureIndices1.w, su, sv, m.coatColor.w), 0.0f, 1.0f);
                    c.sheenColor = reflectanceSpectrum(matRgb(s, m.textureIndices2.w, su, sv, v3(m.sheenColorMode)), wl);
                    c.sheenRoughness = clampf(matScalar(s, m.textureIndices4.y, su, sv, m.transmissionScatter.w), 0.01f, 1.0f);
                    c.sheenZeltner = m.sheenColorMode.w > 0.5f;
                    c.translucentWeight = clampf(matScalar(s, m.textureIndices5.x, su, sv, m.translucentColorWeight.w), 0.0f, 1.0f);
                    c.translucentColor = reflectanceSpectrum(matRgb(s, m.textureIndices5.y, su, sv, v3(m.translucentColorWeight)), wl);

                    // Lobe probabilities (:1435-1464).
                    c.view = normalize(-rd);
                    const float nv = fmaxf(dot(normal, c.view), 0.0f);
                    c.sheenAlbedo = c.sheenZeltner ? sheenAlbedoZeltner(nv, c.sheenRoughness) : sheenAlbedoCK(nv, c.sheenRoughness);
                    c.pSheen = clampf(c.sheen * c.sheenAlbedo * lumGrey(c.sheenColor), 0.0f, 0.95f);
                    c.coatContribution = c.coat > 0.0f ? c.coat * fresnelSchlick(nv, c.coatF0) : 0.0f;
                    c.pCoat = clampf(lumGrey(c.coatContribution), 0.0f, 0.95f);
                    const float baseF = fresnelSchlick(nv, c.f0);
                    c.diffuseContribution = (1.0f - baseF) * c.diffuseWeight * (1.0f - c.translucentWeight) * (1.0f - c.metalness) * c.diffuseColor;
                    c.translucentContribution = (1.0f - baseF) * c.translucentWeight * (1.0f - c.metalness) * c.translucentColor;
                    c.specE = lumGrey(baseF);
                    c.diffE = lumGrey(c.diffuseContribution);
                    c.translucentE = lumGrey(c.translucentContribution);
                    const float baseE = c.diffE + c.translucentE;
                    c.pSpec = clampf(c.specE / fmaxf(c.specE + baseE, 1e-5f), 0.05f, 0.95f);
                    c.pTranslucent = baseE > 1e-5f ? c.translucentE / baseE : 0.0f;
                    const float dSpecE = c.specE + lumGrey(c.coatContribution) + c.pSheen;

                    // Radiance cache: training paths record this vertex (the
                    // radiance after its emission and the throughput reaching
                    // it); other paths stop at terminateBounce and queue a
                    // query for the radiance scattered toward the viewer.
                    bool cacheStop = false;
                    if (cp.enabled) {
                        const std::uint32_t ps = pid / 3U, lane = pid - ps * 3U;
                        RcQuery rq;
                        rq.px = (hp.x - s.boundsLo[0]) * s.boundsInvExtent[0];
                        rq.py = (hp.y - s.boundsLo[1]) * s.boundsInvExtent[1];
                        rq.pz = (hp.z - s.boundsLo[2]) * s.boundsInvExtent[2];
                        rq.dx = c.view.x;
                        rq.dy = c.view.y;
                        rq.dz = c.view.z;
                        rq.lambda0 = pb.lambda0[pid];
                        rq.roughness = 0.5f * (c.rx + c.ry);
                        rq.albedo = clampf(0.2126f * baseRgb.x + 0.7152f * baseRgb.y + 0.0722f * baseRgb.z, 0.0f, 1.0f);
                        rq.lane = lane;
                        rq.path = pid;
                        rq.target = radiance;
                        bool training;
                        std::uint32_t recordBase = 0xffffffffU;
                        if (cp.trainSlot) {
                            const std::uint32_t ts = cp.trainSlot[pid];
                            training = ts != 0xffffffffU;
                            if (training) recordBase = ts * cp.recordVertices;
                        } else {
                            training = (ps + cp.trainOffset) % cp.trainStride == 0U;
                            const std::uint32_t tslot = (ps + cp.trainOffset) / cp.trainStride;
                            if (training && tslot < cp.recordSlots) recordBase = (tslot * 3U + lane) * cp.recordVertices;
                        }
                        if (training) {
                            if (recordBase != 0xffffffffU && static_cast<std::uint32_t>(bounce) < cp.recordVertices) {
                                const std::uint32_t r = recordBase + static_cast<std::uint32_t>(bounce);
                                cp.records[r] = rq;
                                cp.recordThroughput[r] = throughput;
                            }
                        } else {
                            // Path spread at this vertex; a delta lobe before it adds nothing.
                            float2 sp = pb.spread[pid];
                            const float cosIn = fmaxf(fabsf(dot(ogn, rd)), 1e-4f);
                            if (bounce == 0) {
                                sp = make_float2(distance * distance / (4.0f * kPi * cosIn), 0.0f);
                            } else if (!(flags & kPrevDelta) && mis.x > 0.0f) {
                                sp.y += sqrtf(distance * distance / (mis.x * cosIn));
                            }
                            spreadOut = sp;
                            spreadSet = true;
                            // Without the spread rule a path stops exactly at terminateBounce (a path
                            // that scattered in a medium there goes on as a plain path); inference runs
                            // after that bounce only, or after every later one with the rule.
                            if (cp.spreadThreshold <= 0.0f ? bounce == static_cast<int>(cp.terminateBounce)
                                                           : (bounce >= static_cast<int>(cp.terminateBounce) &&
                                                              sp.y * sp.y > cp.spreadThreshold * sp.x)) {
                                cacheStop = true;
                                wantQuery = true;
                                query = rq;
                            }
                        }
                    }

                    float lightTotal = 0.0f;  // set with the light sample; every later use has lightCount > 0
                    if (cacheStop) {
                    } else if (s.lightCount > 0U) {
                        // Analytic NEE (:1469-1592).
                        std::uint32_t li;
                        float selPdf;
                        lightTotal = totalLightWeight(s, hp, normal, c.diffE, dSpecE, c.translucentE);
                        sampleAnalyticLight(s, hp, normal, c.diffE, dSpecE, c.translucentE, lightTotal, rng, li, selPdf);
                        const GpuLightRecord& L = s.lights[li];
                        V3 ld = v3(0.0f, 1.0f, 0.0f);
                        V3 lrad = v3(0.0f);
                        float invPdf = 0.0f, maxShadow = 10000.0f;
                        if (L.textureInfo.z == 0U) {
                            float lp = 0.0f;
                            sampleDomeLight(s, li, L, rng, ld, lp);
                            lrad = sampleDome(s, L, ld);
                            invPdf = lp > 0.0f ? 1.0f / lp : 0.0f;
                        } else if (L.textureInfo.z == 5U) {
                            const float half = 0.5f * (clampf(L.geometry.z, 0.0f, 359.999f) * (kPi / 180.0f));
                            if (half <= 1e-6f) {
                                ld = normalize(v3(L.basisZ));
                            } else {
                                const float ct = mixf(cosf(half), 1.0f, randomFloat(rng));
                                const float stt = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
                                const float phi = kTwoPi * randomFloat(rng);
                                float sp, cp;
                                sincosf(phi, &sp, &cp);
                                ld = normalize(v3(L.basisX) * (cp * stt) + v3(L.basisY) * (sp * stt) + v3(L.basisZ) * ct);
                            }
                            invPdf = 1.0f;
                            lrad = v3(L.colorIntensity) * L.colorIntensity.w;
                        } else {
                            const float lu = randomFloat(rng), lv = randomFloat(rng);
                            V3 lpos = v3(L.position);
                            V3 ln = -normalize(v3(L.basisZ));
                            if (L.textureInfo.z == 1U) {
                                lpos = lpos + v3(L.axisU) * (lu - 0.5f) + v3(L.axisV) * (lv - 0.5f);
                            } else if (L.textureInfo.z == 2U) {
                                const float dr = sqrtf(lu);
                                float sp, cp;
                                sincosf(kTwoPi * lv, &sp, &cp);
                                lpos = lpos + (v3(L.axisU) * cp + v3(L.axisV) * sp) * dr;
                            } else if (L.textureInfo.z == 3U) {
                                ln = uniformSphere(rng);
                                lpos = lpos + ln * L.geometry.x;
                            } else if (L.textureInfo.z == 4U) {
                                float sp, cp;
                                sincosf(kTwoPi * lv, &sp, &cp);
                                ln = normalize(v3(L.basisY) * cp + v3(L.basisZ) * sp);
                                lpos = lpos + v3(L.axisU) * (lu - 0.5f) + ln * L.geometry.x;
                            }
                            const V3 toL = lpos - hp;
                            const float d2 = dot(toL, toL);
                            const float dist = sqrtf(fmaxf(d2, 1e-8f));
                            ld = toL / dist;
                            const float ec = fmaxf(dot(ln, -ld), 0.0f);
                            if (ec > 0.0f && L.controls.w > 0.0f) {
                                const float shapingCos = fmaxf(dot(normalize(v3(L.basisZ)), ld), 0.0f);
                                const float focus = powf(shapingCos, fmaxf(L.shaping.x, 0.0f));
                                const V3 focusColor = v3(mixf(L.focusTint.x, 1.0f, focus), mixf(L.focusTint.y, 1.0f, focus),
                                                         mixf(L.focusTint.z, 1.0f, focus));
                                V3 texv = v3(1.0f);
                                if (L.textureInfo.z == 1U && L.textureInfo.x != kMissingTexture) {
                                    const float4 t = sampleTexture(s, L.textureInfo.x, lu, lv);
                                    texv = v3(t.x, t.y, t.z);
                                }
                                lrad = v3(L.colorIntensity) * L.colorIntensity.w * texv * focusColor * shapingCone(L, shapingCos);
                                invPdf = L.controls.w * ec / fmaxf(d2, 1e-8f);
                                maxShadow = fmaxf(dist - 0.003f, 0.001f);
                            }
                        }
                        if (invPdf > 0.0f) {
                            const float direct = evalDirect(c, ld, L.controls.x, L.controls.y);
                            if (direct > 0.0f) {
                                const float w = L.textureInfo.z == 0U
                                    ? powerHeuristic(environmentLightPdf(s, hp, normal, ld, c.diffE, dSpecE, c.translucentE, lightTotal),
                                                     surfacePdf(c, ld))
                                    : 1.0f;
                                const float contribution = throughput * direct * illuminantSpectrum(lrad, wl) * invPdf * w / fmaxf(selPdf, 1e-8f);
                                queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, ld), ld, maxShadow, contribution, li);
                            }
                        }
                    } else if (!s.hasEmissive) {
                        // Fallback sun and sky (:1593-1668).
                        const V3 up = s.zUp ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
                        const V3 az = s.zUp ? normalize(v3(1.0f, 1.0f, 0.0f)) : normalize(v3(1.0f, 0.0f, 1.0f));
                        const V3 sun = normalize(up * 0.9659258f + az * 0.2588190f);
                        const float direct = evalDirect(c, sun, 1.0f, 1.0f);
                        if (direct > 0.0f) {
                            const float contribution = throughput * direct * illuminantSpectrum(v3(3.45575f, 3.14159f, 2.82743f), wl);
                            queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, sun), sun, 10000.0f, contribution, 0U);
                        }
                        const V3 sky = cosineHemisphere(normal, rng);
                        const float skyDirect = evalDirect(c, sky, 1.0f, 0.0f);
                        if (skyDirect > 0.0f) {
                            const float contribution = throughput * skyDirect * illuminantSpectrum(environmentRgb(s, sky), wl) *
                                                       (kPi / fmaxf(dot(normal, sky), 1e-4f));
                            queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, sky), sky, 10000.0f, contribution, 0U);
                        }
                        if (c.translucentWeight > 0.0f) {
                            const V3 back = cosineHemisphere(-normal, rng);
                            const float backDirect = evalDirect(c, back, 1.0f, 0.0f);
                            if (backDirect > 0.0f) {
                                const float contribution = throughput * backDirect * illuminantSpectrum(environmentRgb(s, back), wl) *
                                                           (kPi / fmaxf(dot(-normal, back), 1e-4f));
                                queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, back), back, 10000.0f, contribution, 0U);
                            }
                        }
                    }
                    if (s.hasEmissive && !cacheStop) {
                        // Emissive-mesh NEE (:1670-1731).
                        const float target = randomFloat(rng);
                        std::uint32_t lo = 0, hi = s.primitiveCount - 1U;
                        while (lo < hi) {
                            const std::uint32_t mid = lo + (hi - lo) / 2U;
                            if (target <= s.emissive[mid].x) hi = mid; else lo = mid + 1U;
                        }
                        const std::uint32_t lprim = lo;
                        const float4 le4 = s.emissive[lprim];
                        if (lprim != prim && le4.y > 0.0f) {
                            const float root = sqrtf(randomFloat(rng));
                            const float second = randomFloat(rng);
                            const float b0 = 1.0f - root, b1 = root * (1.0f - second), b2 = root * second;
                            const std::uint32_t j0 = s.indices[lprim * 3], j1 = s.indices[lprim * 3 + 1], j2 = s.indices[lprim * 3 + 2];
                            const V3 q0 = v3(s.positions[j0]), q1 = v3(s.positions[j1]), q2 = v3(s.positions[j2]);
                            const V3 lpos = q0 * b0 + q1 * b1 + q2 * b2;
                            const V3 ln = normalize(cross(q1 - q0, q2 - q0));
                            const V3 toL = lpos - hp;
                            const float d2 = dot(toL, toL);
                            const float dist = sqrtf(fmaxf(d2, 1e-8f));
                            const V3 ld = toL / dist;
                            const float ec = fabsf(dot(ln, -ld));
                            const float lightPdf = ec <= 0.0f ? 0.0f : le4.y * d2 / ec;
                            if (lightPdf > 0.0f) {
                                const float direct = evalDirect(c, ld, 1.0f, 1.0f);
                                if (direct > 0.0f) {
                                    const GpuMaterialRecord& lm = s.materials[s.triMaterial[lprim]];
                                    const float2 l0 = s.texcoords[lprim * 3], l1 = s.texcoords[lprim * 3 + 1], l2 = s.texcoords[lprim * 3 + 2];
                                    const float lu = l0.x * b0 + l1.x * b1 + l2.x * b2, lvv = l0.y * b0 + l1.y * b1 + l2.y * b2;
                                    const float weight = fmaxf(lm.transmissionOpacityIor.w, 0.0f);
                                    V3 le;
                                    if (lm.textureIndices0.w == kMissingTexture) le = v3(lm.emissionRoughness) * weight;
                                    else {
                                        const float4 e = sampleMaterialTexture(s, lm.textureIndices0.w, lu, lvv, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                                        le = v3(e.x, e.y, e.z) * weight;
                                    }
                                    le = le * materialOpacity<k2L>(s, lprim, b1, b2);
                                    const float w = powerHeuristic(lightPdf, surfacePdf(c, ld));
                                    const float contribution = throughput * direct * illuminantSpectrum(le, wl) * (w / lightPdf);
                                    queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, ld), ld, fmaxf(dist - 0.003f, 0.001f),
                                                contribution, kShadowMeshLight);
                                }
                            }
                        }
                    }

                    if (!lastBounce && !cacheStop) {
                        bool walked = false;
                        if (randomFloat(rng) < c.transmission) {
                            // Transmission (:1735-1777).
                            const float eta = front ? 1.0f / ior : ior;
                            const float cosT = fminf(dot(-rd, normal), 1.0f);
                            float r0 = (1.0f - ior) / (1.0f + ior);
                            r0 *= r0;
                            const float mc = 1.0f - cosT;
                            const float fres = r0 + (1.0f - r0) * powf(mc, 5.0f);
                            const float dni = dot(normal, rd);
                            const float k = 1.0f - eta * eta * (1.0f - dni * dni);
                            const V3 refr = k < 0.0f ? v3(0.0f) : rd * eta - normal * (eta * dni + sqrtf(k));
                            const bool tir = dot(refr, refr) < 1e-8f;
                            const bool thin = m.transmissionColorThinWalled.w > 0.5f;
                            bool through = false;
                            if (tir || randomFloat(rng) < fres) {
                                rd = reflect3(rd, normal);
                            } else {
                                rd = thin ? rd : normalize(refr);
                                through = !thin;
                            }
                            const float depth = fmaxf(m.transmissionMedium.x, 0.0f);
                            throughput *= depth > 0.0f ? 1.0f : transmissionColor;
                            if (through) {
                                if (front) {
                                    const float stc = fmaxf(transmissionColor, 0.001f);
                                    med.x = depth > 0.0f ? -logf(stc) / depth : 0.0f;
                                    const float scat = fmaxf(reconstructRgb(v3(m.transmissionScatter), wl), 0.0f);
                                    med.y = depth > 0.0f ? scat / depth : scat;
                                    med.z = m.transmissionMedium.y;
                                    flags |= kInMedium;
                                } else {
                                    flags &= ~kInMedium;
                                    med.x = 0.0f;
                                    med.y = 0.0f;
                                }
                            }
                            ro = hp + rd * 0.002f;
                            flags |= kEnvOnMiss | kPrevDelta;
                            mis.x = 0.0f;
                            keep = true;
                        } else {
                            // BSDF sampling chain (:1779-1858).
                            flags &= ~kEnvOnMiss;
                            ro = hp + ogn * 0.002f;
                            if (c.pSheen > 0.0f && randomFloat(rng) < c.pSheen) {
                                rd = cosineHemisphere(normal, rng);
                                const float sp = fmaxf(dot(normal, rd), 0.0f) / kPi;
                                throughput *= c.sheen * c.sheenColor * sheenResponse(normal, c.view, rd, c.sheenRoughness, c.sheenZeltner) /
                                              fmaxf(c.pSheen * sp, 1e-6f);
                                flags |= kEnvOnMiss;
                            } else {
                                throughput *= (1.0f - c.sheen * c.sheenAlbedo) / fmaxf(1.0f - c.pSheen, 1e-4f);
                                if (c.pCoat > 0.0f && randomFloat(rng) < c.pCoat) {
                                    const V3 inc = rd;
                                    rd = sampleGgx(rd, normal, c.tangent, c.crx, c.cry, rng);
                                    throughput *= c.coat * ggxSampleWeight(inc, rd, normal, c.tangent, c.crx, c.cry, c.coatF0) / c.pCoat;
                                    if (fmaxf(c.crx, c.cry) < 0.2f) flags |= kEnvOnMiss;
                                } else {
                                    throughput *= (1.0f - c.coatContribution) / fmaxf(1.0f - c.pCoat, 1e-4f);
                                    if (randomFloat(rng) < c.pSpec) {
                                        const V3 inc = rd;
                                        rd = sampleGgx(rd, normal, c.tangent, c.rx, c.ry, rng);
                                        throughput *= ggxSampleWeight(inc, rd, normal, c.tangent, c.rx, c.ry, c.f0) / c.pSpec;
                                        if (fmaxf(c.rx, c.ry) < 0.2f) flags |= kEnvOnMiss;
                                    } else if (c.pTranslucent > 0.0f && randomFloat(rng) < c.pTranslucent) {
                                        throughput *= c.translucentContribution / fmaxf((1.0f - c.pSpec) * c.pTranslucent, 1e-4f);
                                        rd = cosineHemisphere(-normal, rng);
                                        ro = offsetOrigin(hp, ogn, rd);
                                    } else {
                                        throughput *= c.diffuseContribution / fmaxf((1.0f - c.pSpec) * (1.0f - c.pTranslucent), 1e-4f);
                                        if (c.subsurface > 0.0f && randomFloat(rng) < c.subsurface) {
                                            const float radius = fmaxf(reconstructRgb(sssRadiusRgb, wl) * fmaxf(m.subsurfaceWeightScale.y, 0.0f), 0.001f);
                                            const float albedo = reflectanceSpectrum(sssColorRgb, wl);
                                            V3 eo, ed;
                                            float att = 1.0f;
                                            bool exited = false;
                                            if constexpr (kWalk == kShadeWalkDefer) {
                                                deferred = true;
                                                walkIn0 = make_float4(hp.x, hp.y, hp.z, radius);
                                                walkIn1 = make_float4(ogn.x, ogn.y, ogn.z, albedo);
                                                walkIn2 = make_float4(m.subsurfaceWeightScale.z, __uint_as_float(rng), 0.0f, 0.0f);
                                            } else if constexpr (kWalk == kShadeWalkResume) {
                                                const float4 w0 = pb.walk[3U * i], w1 = pb.walk[3U * i + 1U];
                                                rng = __float_as_uint(pb.walk[3U * i + 2U].y);
                                                exited = w1.w != 0.0f;
                                                eo = v3(w0);
                                                ed = v3(w1);
                                                att = w0.w;
                                            } else {
                                                exited = randomWalk<k2L>(s, hp, ogn, radius, albedo, m.subsurfaceWeightScale.z, rng, st, eo, ed, att);
                                            }
                                            if (exited) {
                                                throughput *= att;
                                                ro = eo;
                                                rd = ed;
                                                flags |= kEnvOnMiss | kPrevDelta;
                                                mis.x = 0.0f;
                                                walked = true;
                                            }
                                        }
                                        if (!walked && !deferred) {
                                            rd = cosineHemisphere(normal, rng);
                                            const float lambert = fmaxf(dot(normal, rd), 0.0f) / kPi;
                                            throughput *= evalDiffuse(normal, c.view, rd, c.diffuseColor, c.diffuseRoughness, c.diffuseModel) /
                                                          fmaxf(lambert, 1e-6f);
                                        }
                                    }
                                }
                            }
                            if (walked) {
                                keep = true;
                            } else if (!deferred) {
                                if (s.lightCount > 0U || s.hasEmissive) {
                                    mis.x = surfacePdf(c, rd);
                                    flags &= ~kPrevDelta;
                                    flags |= kEnvOnMiss;
                                    mis.y = s.lightCount > 0U
                                        ? environmentLightPdf(s, hp, normal, rd, c.diffE, dSpecE, c.translucentE, lightTotal)
                                        : 0.0f;
                                }
                                keep = true;
                                // Russian roulette (:1875-1879).
                                if (bounce >= 2) {
                                    const float survival = clampf(throughput, 0.1f, 0.95f);
                                    if (randomFloat(rng) > survival) keep = false;
                                    else throughput /= survival;
                                }
                            }
                        }
                    }
                }
                if (!deferred) {
                    if (keep) {
                        pb.rayO[pid] = make_float4(ro.x, ro.y, ro.z, 0.0f);
                        pb.rayD[pid] = make_float4(rd.x, rd.y, rd.z, 0.0f);
                        pb.flags[pid] = flags;
                        pb.mis[pid] = mis;
                        pb.medium[pid] = med;
                    }
                    pb.rng[pid] = rng;
                    pb.throughput[pid] = throughput;
                    pb.radiance[pid] = radiance;
                    if (spreadSet) pb.spread[pid] = spreadOut;
                }
            }
        }
        waveAppend(keep, pid, outQueue, outCounter);
        if constexpr (kWalk != kShadeWalkResume) {
            // A resumed path queued its shadow rays when it was deferred.
            waveAppend(slot > 0, pid, pb.shadowQueue, pb.counters + 2 * kCounterStride);
            if (cp.enabled) {
                const std::uint32_t qslot = waveSlot(wantQuery, cp.queryCount);
                if (wantQuery && qslot < cp.queryCapacity) cp.queries[qslot] = query;
            }
        }
        if constexpr (kWalk == kShadeWalkDefer) {
            const std::uint32_t wslot = waveSlot(deferred, pb.counters + 3 * kCounterStride);
            if (deferred) {
                pb.walk[3U * wslot] = walkIn0;
                pb.walk[3U * wslot + 1U] = walkIn1;
                pb.walk[3U * wslot + 2U] = walkIn2;
                pb.walkQueue[wslot] = pid;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// walk: the subsurface random walks that shade deferred, one per lane.
// ---------------------------------------------------------------------------
template <class Scenes, bool k2L>
__global__ __launch_bounds__(kBlock) void walkKernel(Scenes scenes, PathBuffers pb)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
    const std::uint32_t count = pb.counters[3 * kCounterStride];
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const DeviceScene& s = scenes.Get(pb.walkQueue[i]);
        const float4 w0 = pb.walk[3U * i], w1 = pb.walk[3U * i + 1U], w2 = pb.walk[3U * i + 2U];
        std::uint32_t rng = __float_as_uint(w2.y);
        V3 eo = v3(0.0f), ed = v3(0.0f);
        float att = 1.0f;
        const bool exited = randomWalk<k2L>(s, v3(w0), v3(w1), w0.w, w1.w, w2.x, rng, st, eo, ed, att);
        pb.walk[3U * i] = make_float4(eo.x, eo.y, eo.z, att);
        pb.walk[3U * i + 1U] = make_float4(ed.x, ed.y, ed.z, exited ? 1.0f : 0.0f);
        pb.walk[3U * i + 2U] = make_float4(w2.x, __uint_as_float(rng), 0.0f, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// shadow: visibility of each queued path's shadow rays, added to its
// radiance in slot order (shadowVisibility :855-889, meshLightVisible
// :891-912).
// ---------------------------------------------------------------------------
template <bool kCount, class Scenes, bool k2L>
__global__ __launch_bounds__(kBlock) void shadowKernel(Scenes scenes, PathBuffers pb, bool shared, TraversalCounters* ctr)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
    __shared__ unsigned int tracedInBlock;
    if (threadIdx.x == 0) tracedInBlock = 0U;
    __syncthreads();
    unsigned int traced = 0U;
    const std::uint32_t queued = pb.counters[2 * kCounterStride];
    if (blockIdx.x == 0 && threadIdx.x == 0) pb.totals[1] += queued;
    // shared: shadowPairKernel grouped the lanes of a pixel sample with
    // equal shadow rays; each group's rays are traced once and the
    // visibility applied to every lane of the group.
    const std::uint32_t count = shared ? pb.counters[6 * kCounterStride] : queued;
    const std::uint32_t* queue = shared ? pb.shadowLeaders : pb.shadowQueue;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const std::uint32_t pid = queue[i];
        const DeviceScene& s = scenes.Get(pid);
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
        for (int k = 0; k < kShadowSlots; ++k) {
            const std::size_t idx = static_cast<std::size_t>(k) * pb.capacity + pid;
            const std::uint32_t light = pb.shadowLight[idx];
            if (light == 0xffffffffU) break;
            const float4 o = pb.shadowO[idx], d = pb.shadowD[idx];
            float tmax = o.w;
            const bool mesh = light == kShadowMeshLight;
            bool trace = true;
            int outcome = 0;  // 1: occluded
            float occluderDistance = 0.0f;
            bool falloffSet = false;
            float4 shadow = make_float4(1.0f, -1.0f, -1.0f, 1.0f);
            float4 shadowColor = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (!mesh) {
                const GpuLightRecord& L = s.lights[light];
                shadow = make_float4(L.shadow.x, L.shadow.y, L.shadow.z, L.shadow.w);
                shadowColor = make_float4(L.shadowColor.x, L.shadowColor.y, L.shadowColor.z, 0.0f);
                if (shadow.x < 0.5f) trace = false;
                if (shadow.y >= 0.0f) tmax = fminf(tmax, shadow.y);
            }
            if (tmax <= 0.001f) trace = false;
            if (trace) {
                ++traced;
                const RayPre r = prepareRay(o.x, o.y, o.z, d.x, d.y, d.z);
                const bool falloff = !mesh && shadow.y >= 0.0f && shadow.z > 0.0f;
                falloffSet = falloff;
                bool occluded;
                float hitDistance = 0.0f;
                if (s.opaque && !falloff) {
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
                        if (s.opaque || materialOpacity<k2L>(s, h.prim, h.u, h.v) >= 0.5f) {
                            occluded = true;
                            hitDistance = h.t;
                            break;
                        }
                        loT = h.t;
                        loPrim = h.prim;
                    }
                }
                outcome = occluded ? 1 : 0;
                occluderDistance = hitDistance;
            }
            // Visibility of each lane of the group with its own wavelength, and
            // its own contribution, in slot order.
            for (int m = 0; m < members; ++m) {
                const float wl = wls[m];
                const float contribution = m == 0 ? d.w : pb.shadowD[static_cast<std::size_t>(k) * pb.capacity + member[m]].w;
                float visibility = 1.0f;
                if (trace && outcome == 1) {
                    if (mesh) {
                        visibility = 0.0f;
                    } else {
                        float strength = 1.0f;
                        if (falloffSet) {
                            const float f = fminf(shadow.z, fmaxf(shadow.y, 0.0f));
                            strength = clampf((shadow.y - occluderDistance) / fmaxf(f, 1e-6f), 0.0f, 1.0f);
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
__global__ void gatherTrainingKernel(PathBuffers pb, CacheParams cp, std::uint32_t records)
{
    if (cp.recordCount) records = min(*cp.recordCount, cp.recordSlots) * cp.recordVertices;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < records; i += gridDim.x * blockDim.x) {
        RcQuery r = cp.records[i];
        if (r.lane > 2U) continue;
        const float t = cp.recordThroughput[i];
        if (!(t > 0.0f)) continue;
        const float target = (pb.radiance[r.path] - r.target) / t;
        if (!isfinite(target)) continue;
        r.target = fmaxf(target, 0.0f);
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
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
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
    TraversalStack st;
    st.lds = stackMem + threadIdx.x;
    st.stride = kBlock;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        const RayPre r = prepareRay(a.x, a.y, a.z, b.x, b.y, b.z);
        out[i] = traceAnyS<false, k2L>(s, r, a.w, b.w, st, nullptr) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static unsigned GridFor(std::uint32_t work, unsigned maxBlocks)
{
    const unsigned blocks = (work + kBlock - 1U) / kBlock;
    return blocks == 0 ? 1U : (blocks < maxBlocks ? blocks : maxBlocks);
}

hipError_t LaunchRaygen(const PathBuffers& pb, const ChunkDesc& c, std::uint32_t paths, const CacheParams& cp,
                        unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(raygenKernel, dim3(GridFor(paths, maxBlocks)), dim3(kBlock), 0, st, pb, c, paths, cp);
    return hipGetLastError();
}
hipError_t LaunchGatherTraining(const PathBuffers& pb, const CacheParams& cp, std::uint32_t records, unsigned maxBlocks,
                                hipStream_t st, const float* outputs, const float* calibration)
{
    hipLaunchKernelGGL(gatherTrainingKernel, dim3(GridFor(records, maxBlocks)), dim3(kBlock), 0, st, pb, cp, records);
    hipLaunchKernelGGL(finishTrainingKernel, dim3(GridFor(records, maxBlocks)), dim3(kBlock), 0, st, pb, cp, records, outputs,
                       calibration);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchExtendT(const Scenes& s, const PathBuffers& pb, std::uint32_t in, int sharing, std::uint32_t maxWork,
                         unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    if (ctr)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(extendKernel<true, Scenes, k2L>), dim3(GridFor(maxWork, maxBlocks)), dim3(kBlock),
                           0, st, s, pb, in, sharing, ctr);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(extendKernel<false, Scenes, k2L>), dim3(GridFor(maxWork, maxBlocks)),
                           dim3(kBlock), 0, st, s, pb, in, sharing, ctr);
    return hipGetLastError();
}
template <class Scenes>
hipError_t LaunchPairT(const Scenes& s, const PathBuffers& pb, std::uint32_t out, std::uint32_t maxWork,
                       unsigned maxBlocks, hipStream_t st)
{
    const unsigned blocks = std::max(1U, std::min((maxWork + kPairBlock - 1U) / kPairBlock, maxBlocks));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(pairKernel<Scenes>), dim3(blocks), dim3(kPairBlock), 0, st, s, pb, out);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchShadeT(const Scenes& s, const PathBuffers& pb, std::uint32_t in, int bounce, int last,
                        const CacheParams& cp, int walk, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    const dim3 grid(GridFor(maxWork, maxBlocks));
    if (walk == kShadeWalkDefer)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkDefer>), grid, dim3(kBlock), 0, st, s, pb, in,
                           bounce, last, cp);
    else if (walk == kShadeWalkResume)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkResume>), grid, dim3(kBlock), 0, st, s, pb, in,
                           bounce, last, cp);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkInline>), grid, dim3(kBlock), 0, st, s, pb, in,
                           bounce, last, cp);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchWalkT(const Scenes& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(HIP_KERNEL_NAME(walkKernel<Scenes, k2L>), dim3(GridFor(maxWork, maxBlocks)), dim3(kBlock), 0, st, s,
                       pb);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchShadowT(const Scenes& s, const PathBuffers& pb, bool shared, std::uint32_t maxWork, unsigned maxBlocks,
                         TraversalCounters* ctr, hipStream_t st)
{
    if (ctr)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowKernel<true, Scenes, k2L>), dim3(GridFor(maxWork, maxBlocks)), dim3(kBlock),
                           0, st, s, pb, shared, ctr);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowKernel<false, Scenes, k2L>), dim3(GridFor(maxWork, maxBlocks)),
                           dim3(kBlock), 0, st, s, pb, shared, ctr);
    return hipGetLastError();
}
hipError_t LaunchExtend(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int sharing, std::uint32_t maxWork,
                        unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    return s.instanceCount ? LaunchExtendT<SingleScene, true>(SingleScene{s}, pb, in, sharing, maxWork, maxBlocks, ctr, st)
                           : LaunchExtendT<SingleScene, false>(SingleScene{s}, pb, in, sharing, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchPair(const DeviceScene& s, const PathBuffers& pb, std::uint32_t out, std::uint32_t maxWork,
                      unsigned maxBlocks, hipStream_t st)
{
    return LaunchPairT<SingleScene>(SingleScene{s}, pb, out, maxWork, maxBlocks, st);
}
hipError_t LaunchShade(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int bounce, int last,
                       const CacheParams& cp, int walk, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    return s.instanceCount
        ? LaunchShadeT<SingleScene, true>(SingleScene{s}, pb, in, bounce, last, cp, walk, maxWork, maxBlocks, st)
        : LaunchShadeT<SingleScene, false>(SingleScene{s}, pb, in, bounce, last, cp, walk, maxWork, maxBlocks, st);
}
hipError_t LaunchWalk(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                      hipStream_t st)
{
    return s.instanceCount ? LaunchWalkT<SingleScene, true>(SingleScene{s}, pb, maxWork, maxBlocks, st)
                           : LaunchWalkT<SingleScene, false>(SingleScene{s}, pb, maxWork, maxBlocks, st);
}
hipError_t LaunchShadowPair(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st)
{
    const unsigned blocks = std::max(1U, std::min((maxWork + kPairBlock - 1U) / kPairBlock, maxBlocks));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowPairKernel<SingleScene>), dim3(blocks), dim3(kPairBlock), 0, st,
                       SingleScene{s}, pb);
    return hipGetLastError();
}
hipError_t LaunchShadow(const DeviceScene& s, const PathBuffers& pb, bool shared, std::uint32_t maxWork, unsigned maxBlocks,
                        TraversalCounters* ctr, hipStream_t st)
{
    return s.instanceCount ? LaunchShadowT<SingleScene, true>(SingleScene{s}, pb, shared, maxWork, maxBlocks, ctr, st)
                           : LaunchShadowT<SingleScene, false>(SingleScene{s}, pb, shared, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchExtendBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t in,
                             int sharing, std::uint32_t maxWork, unsigned maxBlocks, TraversalCounters* ctr,
                             hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchExtendT<SceneTable, true>(t, pb, in, sharing, maxWork, maxBlocks, ctr, st)
                    : LaunchExtendT<SceneTable, false>(t, pb, in, sharing, maxWork, maxBlocks, ctr, st);
}
hipError_t LaunchPairBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t out, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st)
{
    return LaunchPairT<SceneTable>(SceneTable{table, pb.pathScene}, pb, out, maxWork, maxBlocks, st);
}
hipError_t LaunchShadeBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t in, int bounce,
                            int last, const CacheParams& cp, int walk, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchShadeT<SceneTable, true>(t, pb, in, bounce, last, cp, walk, maxWork, maxBlocks, st)
                    : LaunchShadeT<SceneTable, false>(t, pb, in, bounce, last, cp, walk, maxWork, maxBlocks, st);
}
hipError_t LaunchWalkBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchWalkT<SceneTable, true>(t, pb, maxWork, maxBlocks, st)
                    : LaunchWalkT<SceneTable, false>(t, pb, maxWork, maxBlocks, st);
}
hipError_t LaunchShadowPairBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t maxWork,
                                 unsigned maxBlocks, hipStream_t st)
{
    const unsigned blocks = std::max(1U, std::min((maxWork + kPairBlock - 1U) / kPairBlock, maxBlocks));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(shadowPairKernel<SceneTable>), dim3(blocks), dim3(kPairBlock), 0, st,
                       SceneTable{table, pb.pathScene}, pb);
    return hipGetLastError();
}
hipError_t LaunchShadowBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t maxWork,
                             unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchShadowT<SceneTable, true>(t, pb, true, maxWork, maxBlocks, ctr, st)
                    : LaunchShadowT<SceneTable, false>(t, pb, true, maxWork, maxBlocks, ctr, st);
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
