// Wavefront building blocks shared by the stage kernels (kernels.hip,
// shade.hip): block size, queue appends, the opacity of a hit.
#pragma once

#include "dev_shading.h"
#include "dev_traverse2.h"

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

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

constexpr unsigned kPairBlock = 1024;


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

// pairAppend for threads that each hold three consecutive path ids
// (pid0 .. pid0 + 2), appended in thread order.
__device__ __forceinline__ void pairAppend3(bool want, std::uint32_t pid0, std::uint32_t* queue, std::uint32_t* counter,
                                            std::uint32_t* waveCount, std::uint32_t* waveBase)
{
    constexpr unsigned kWaves = kPairBlock / 64U;
    const std::uint32_t lane = laneId(), wave = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(want);
    if (lane == 0U) waveCount[wave] = 3U * static_cast<std::uint32_t>(__popcll(mask));
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
    if (want) {
        const std::uint32_t at = waveBase[wave] + 3U * static_cast<std::uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
        queue[at] = pid0;
        queue[at + 1U] = pid0 + 1U;
        queue[at + 2U] = pid0 + 2U;
    }
    __syncthreads();
}

// Leader / follower decision of one path among the lanes of its pixel
// sample, from kWords words that must be equal bit for bit (the lane
// pairing at the end of shade). Returns the follower mask of a path that
// traces (bit j: pid + 1 + j) in `followers`; true when the path traces.
// (Skipping the comparisons of words or distances that no lane of the wave
// needs, behind ballots, measured slower.)
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

// Blocks for `work` items, at most maxBlocks (the stage loops are
// grid-stride).
inline unsigned GridFor(std::uint32_t work, unsigned maxBlocks)
{
    const unsigned blocks = (work + kBlock - 1U) / kBlock;
    return blocks == 0 ? 1U : (blocks < maxBlocks ? blocks : maxBlocks);
}

// Blocks of `kernel` (blockSize threads, no dynamic LDS) the current device
// keeps resident at once: its occupancy limit per CU times the CU count,
// queried once per kernel and device.
inline unsigned ResidentBlocks(const void* kernel, int blockSize)
{
    static std::mutex mutex;
    static std::map<std::pair<const void*, int>, unsigned> cache;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return 0U;
    const std::lock_guard<std::mutex> lock(mutex);
    const auto key = std::make_pair(kernel, device);
    if (const auto found = cache.find(key); found != cache.end()) return found->second;
    int perCu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, blockSize, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0U;
    const unsigned resident = static_cast<unsigned>(perCu > 0 ? perCu : 1) * static_cast<unsigned>(cus > 0 ? cus : 1);
    cache.emplace(key, resident);
    return resident;
}

// Grid of a grid-stride stage kernel: blocks for `work` items, at most
// maxBlocks and at most what the device keeps resident. Every wave of such
// a loop does about the same work, so blocks beyond the resident ones run
// as a partial second round: at 6 resident waves per SIMD and 8 waves of
// blocks, a third of the GPU would idle for a whole per-wave loop at the
// end of each launch.
template <class Kernel>
unsigned StageGrid(Kernel kernel, std::uint32_t work, unsigned maxBlocks, int blockSize = kBlock)
{
    const std::uint32_t blocks = (work + static_cast<std::uint32_t>(blockSize) - 1U) / static_cast<std::uint32_t>(blockSize);
    unsigned grid = blocks == 0 ? 1U : (blocks < maxBlocks ? blocks : maxBlocks);
    const unsigned resident = ResidentBlocks(reinterpret_cast<const void*>(kernel), blockSize);
    if (resident > 0U && grid > resident) grid = resident;
    return grid;
}

} // namespace hdcx
