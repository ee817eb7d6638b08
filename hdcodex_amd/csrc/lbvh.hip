// Device BVH build for gfx950: a Morton-ordered binary radix tree (Karras,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees", HPG 2012) under binned-SAH top levels (the HLBVH split of
// Garanzha, Pantaleoni and McAllister, HPG 2011), collapsed to the four-wide
// node format the traversal uses; and a bottom-up refit for moved vertices.
// The reference leaves both to the Vulkan driver
// (src/gpu/vulkan_path_tracer.cpp:3032-3125).
//
//   1. Morton code (10 bits per axis) of each triangle's box centre, sorted
//      with its triangle id by a stable LSD radix sort (30 key bits, three
//      passes of 10 bits; below).
//   2. Binary radix tree, one thread per inner node.
//   3. Boxes bottom-up: the second child to arrive at a node unions both
//      children (per-node counter, agent-scope acquire/release).
//   4. Top levels: the frontier is every node of <= S triangles whose parent
//      has more (S = n / targetClusters); the host builds a binned-SAH tree
//      over the frontier boxes (a few thousand) and appends it above them.
//   5. Collapse to BVH4 one level per launch; node ids come from an atomic
//      counter, so each level is one contiguous id range (the refit order).
//   6. Triangle slots in sorted order.
// The per-element steps live in lbvh_core.h, shared with a serial CPU run.
#include "lbvh.h"
#include "lbvh_core.h"
#include "lbvh_host.hpp"

#include <algorithm>

namespace hdcx {
namespace {

using lbvh::Box4;
using lbvh::FrontierItem;

constexpr unsigned kBlock = 256;

unsigned Grid(std::uint32_t n) { return std::max(1U, std::min((n + kBlock - 1) / kBlock, 16384U)); }

struct DeviceAtomics {
    __device__ static std::uint32_t Arrive(std::uint32_t* c)
    {
        return __hip_atomic_fetch_add(c, 1U, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static std::uint32_t Append(std::uint32_t* c) { return atomicAdd(c, 1U); }
    __device__ static void Max(std::uint32_t* c, std::uint32_t v) { atomicMax(c, v); }
    __device__ static void Flag(std::uint32_t* c) { atomicOr(c, 1U); }
};

#define HDCX_GRID_STRIDE(i, begin, end) \
    for (std::uint32_t i = (begin) + blockIdx.x * blockDim.x + threadIdx.x; i < (end); i += gridDim.x * blockDim.x)

// ---- Stable LSD radix sort of (key, value) pairs, 10-bit digits ----
//
// One wave per tile of kSortTile consecutive pairs. A pass over digit
// (key >> shift) & 1023:
//   histogram: each tile counts its digits (LDS) into hist[tile * 1024 + d];
//   scan (one block of 1024 threads, thread d): hist[t][d] becomes the count
//     of digit d in tiles before t, and start[d] the count of all digits
//     below d (an exclusive scan of the digit totals);
//   scatter: each tile walks its pairs in order, 64 per round; the lanes of
//     a round with equal digits find each other with ten ballots (one per
//     digit bit), take consecutive places in lane order after the digit's
//     running count in LDS, and write to start[d] + hist[t][d] + place.
// Pairs of equal digit keep their order within a round, across rounds and
// across tiles, so the sort is stable (the serial run of the builder uses
// std::stable_sort on the same keys).
constexpr std::uint32_t kSortDigits = 1024;
constexpr std::uint32_t kSortTile = 64U * 64U;

__global__ __launch_bounds__(64) void sortHistogramKernel(const std::uint32_t* __restrict__ keys, std::uint32_t n,
                                                          std::uint32_t shift, std::uint32_t* __restrict__ hist)
{
    __shared__ std::uint32_t count[kSortDigits];
    for (std::uint32_t d = threadIdx.x; d < kSortDigits; d += 64U) count[d] = 0U;
    __syncthreads();
    const std::uint32_t base = blockIdx.x * kSortTile;
    const std::uint32_t end = min(n, base + kSortTile);
    for (std::uint32_t i = base + threadIdx.x; i < end; i += 64U) atomicAdd(&count[(keys[i] >> shift) & (kSortDigits - 1U)], 1U);
    __syncthreads();
    for (std::uint32_t d = threadIdx.x; d < kSortDigits; d += 64U) hist[static_cast<std::size_t>(blockIdx.x) * kSortDigits + d] = count[d];
}

__global__ __launch_bounds__(kSortDigits) void sortScanKernel(std::uint32_t* __restrict__ hist, std::uint32_t tiles,
                                                              std::uint32_t* __restrict__ start)
{
    __shared__ std::uint32_t total[kSortDigits];
    const std::uint32_t d = threadIdx.x;
    std::uint32_t running = 0U;
    for (std::uint32_t t = 0; t < tiles; ++t) {
        const std::size_t k = static_cast<std::size_t>(t) * kSortDigits + d;
        const std::uint32_t c = hist[k];
        hist[k] = running;
        running += c;
    }
    total[d] = running;
    __syncthreads();
    // Exclusive scan of the 1024 totals (Hillis-Steele on an inclusive copy).
    for (std::uint32_t off = 1U; off < kSortDigits; off <<= 1) {
        const std::uint32_t v = d >= off ? total[d - off] : 0U;
        __syncthreads();
        total[d] += v;
        __syncthreads();
    }
    start[d] = total[d] - running;
}

__global__ __launch_bounds__(64) void sortScatterKernel(const std::uint32_t* __restrict__ keys,
                                                        const std::uint32_t* __restrict__ values, std::uint32_t n,
                                                        std::uint32_t shift, const std::uint32_t* __restrict__ hist,
                                                        const std::uint32_t* __restrict__ start,
                                                        std::uint32_t* __restrict__ keysOut,
                                                        std::uint32_t* __restrict__ valuesOut)
{
    __shared__ std::uint32_t next[kSortDigits];  // output index of the next pair of each digit
    const std::size_t row = static_cast<std::size_t>(blockIdx.x) * kSortDigits;
    for (std::uint32_t d = threadIdx.x; d < kSortDigits; d += 64U) next[d] = start[d] + hist[row + d];
    __syncthreads();
    const std::uint32_t lane = threadIdx.x;
    const unsigned long long below = (1ULL << lane) - 1ULL;
    const std::uint32_t base = blockIdx.x * kSortTile;
    const std::uint32_t end = min(n, base + kSortTile);
    for (std::uint32_t r = base; r < end; r += 64U) {
        const std::uint32_t i = r + lane;
        const bool valid = i < end;
        const std::uint32_t key = valid ? keys[i] : 0U;
        const std::uint32_t value = valid ? values[i] : 0U;
        const std::uint32_t digit = (key >> shift) & (kSortDigits - 1U);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (std::uint32_t b = 0; b < 10U; ++b) {
            const bool set = (digit >> b) & 1U;
            const unsigned long long m = __ballot(set);
            peers &= set ? m : ~m;
        }
        // Lanes with this lane's digit (itself included, when valid); the
        // lowest of them reserves places for the whole group.
        const std::uint32_t place = static_cast<std::uint32_t>(__popcll(peers & below));
        const int first = __ffsll(static_cast<long long>(peers)) - 1;
        std::uint32_t at = 0U;
        if (valid && place == 0U) {
            at = next[digit];
            next[digit] = at + static_cast<std::uint32_t>(__popcll(peers));
        }
        at = __shfl(at, valid ? first : 0);
        if (valid) {
            keysOut[at + place] = key;
            valuesOut[at + place] = value;
        }
    }
}

__global__ void prepareKernel(const float* __restrict__ pos, const std::uint32_t* __restrict__ idx, std::uint32_t n,
                              float lox, float loy, float loz, float sx, float sy, float sz,
                              std::uint32_t* __restrict__ morton, std::uint32_t* __restrict__ ids,
                              Box4* __restrict__ triBox)
{
    const float lo[3] = {lox, loy, loz}, scale[3] = {sx, sy, sz};
    HDCX_GRID_STRIDE(i, 0U, n)
    {
        Box4 b;
        morton[i] = lbvh::PrepareTriangle(pos, idx, i, lo, scale, b);
        triBox[i] = b;
        ids[i] = i;
    }
}

__global__ void radixTreeKernel(const std::uint32_t* __restrict__ keys, std::uint32_t n, std::uint32_t* child2,
                                std::uint32_t* parent, std::uint32_t* leafParent, std::uint32_t* range2)
{
    HDCX_GRID_STRIDE(i, 0U, n - 1U)
    lbvh::RadixTreeNode(keys, static_cast<int>(n), static_cast<int>(i), child2, parent, leafParent, range2);
}

__global__ void bottomUpKernel(std::uint32_t n, const std::uint32_t* child2, const std::uint32_t* parent,
                               const std::uint32_t* leafParent, std::uint32_t* arrivals, const std::uint32_t* ids,
                               const Box4* triBox, const std::uint32_t* range2, Box4* nodeBox, float* nodeCost,
                               std::uint32_t* leafFlag)
{
    HDCX_GRID_STRIDE(k, 0U, n)
    lbvh::BottomUpLeaf<DeviceAtomics>(k, child2, parent, leafParent, arrivals, ids, triBox, range2, nodeBox, nodeCost,
                                      leafFlag);
}

__global__ void frontierKernel(std::uint32_t n, std::uint32_t S, float maxArea, const std::uint32_t* child2,
                               const std::uint32_t* range2, const Box4* nodeBox, const Box4* triBox,
                               const std::uint32_t* ids, std::uint32_t* count, std::uint32_t capacity, FrontierItem* out)
{
    HDCX_GRID_STRIDE(i, 0U, n - 1U)
    lbvh::FrontierNode<DeviceAtomics>(i, S, maxArea, child2, range2, nodeBox, triBox, ids, count, capacity, out);
}

__global__ void collapseKernel(std::uint32_t begin, std::uint32_t end, std::uint32_t level, std::uint32_t* source,
                               const std::uint32_t* child2, const std::uint32_t* range2, const Box4* nodeBox,
                               const Box4* triBox, const std::uint32_t* ids, const std::uint32_t* leafFlag,
                               std::uint32_t* nodeStack, std::uint32_t* counters, std::uint32_t capacity,
                               Bvh4Node* nodes)
{
    HDCX_GRID_STRIDE(id, begin, end)
    lbvh::CollapseNode<DeviceAtomics>(id, level, source, child2, range2, nodeBox, triBox, ids, leafFlag, nodeStack,
                                      counters, capacity, nodes);
}

__global__ void slotsKernel(std::uint32_t n, const std::uint32_t* __restrict__ ids, const float* __restrict__ pos,
                            const std::uint32_t* __restrict__ idx, float* __restrict__ slots)
{
    HDCX_GRID_STRIDE(k, 0U, n)
    lbvh::WriteSlot(pos, idx, ids[k], slots + 12 * static_cast<std::size_t>(k));
}

__global__ void refitSlotsKernel(std::uint32_t count, const float* __restrict__ pos,
                                 const std::uint32_t* __restrict__ idx, float* __restrict__ slots)
{
    HDCX_GRID_STRIDE(k, 0U, count)
    {
        float* s = slots + 12 * static_cast<std::size_t>(k);
        lbvh::WriteSlot(pos, idx, lbvh::FloatToBits(s[3]), s);
    }
}

__global__ void refitLevelKernel(const std::uint32_t* __restrict__ order, std::uint32_t begin, std::uint32_t end,
                                 Bvh4Node* nodes, const float* __restrict__ slots)
{
    HDCX_GRID_STRIDE(w, begin, end)
    lbvh::RefitNode(nodes[order ? order[w] : w], nodes, slots);
}

#define HDCX_LAUNCH(kernel, grid, ...)                                             \
    do {                                                                           \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, st, __VA_ARGS__); \
        Check(hipGetLastError(), #kernel);                                         \
    } while (0)

} // namespace

// Sorts n (key, value) pairs by the low 30 key bits, stably: three passes
// of 10 bits, (a, b) -> (sa, sb) -> (a, b) -> (sa, sb).
void DeviceBvhBuilder::SortPairs30(std::uint32_t* a, std::uint32_t* b, std::uint32_t* sa, std::uint32_t* sb,
                                   std::uint32_t n, hipStream_t st)
{
    const std::uint32_t tiles = (n + kSortTile - 1U) / kSortTile;
    temp_.Reserve(sizeof(std::uint32_t) * (static_cast<std::size_t>(tiles) * kSortDigits + kSortDigits));
    std::uint32_t* hist = temp_.as<std::uint32_t>();
    std::uint32_t* start = hist + static_cast<std::size_t>(tiles) * kSortDigits;
    std::uint32_t* inK = a;
    std::uint32_t* inV = b;
    std::uint32_t* outK = sa;
    std::uint32_t* outV = sb;
    for (std::uint32_t shift = 0; shift < 30U; shift += 10U) {
        hipLaunchKernelGGL(sortHistogramKernel, dim3(tiles), dim3(64), 0, st, inK, n, shift, hist);
        Check(hipGetLastError(), "sort histogram");
        hipLaunchKernelGGL(sortScanKernel, dim3(1), dim3(kSortDigits), 0, st, hist, tiles, start);
        Check(hipGetLastError(), "sort scan");
        hipLaunchKernelGGL(sortScatterKernel, dim3(tiles), dim3(64), 0, st, inK, inV, n, shift, hist, start, outK, outV);
        Check(hipGetLastError(), "sort scatter");
        std::swap(inK, outK);
        std::swap(inV, outV);
    }
}

void DeviceBvhBuilder::Build(const float4* positions, const std::uint32_t* indices, std::uint32_t n, const float lo[3],
                             const float hi[3], std::uint32_t targetClusters, hipStream_t st, DeviceBuffer& nodes,
                             DeviceBuffer& slots, RefitPlan& plan, DeviceBvhStats& stats)
{
    if (n < 2 || n > (1U << 28)) throw std::invalid_argument("device BVH build needs 2 .. 2^28 triangles");
    stats = DeviceBvhStats{};
    const std::uint32_t S = FrontierThreshold(n, targetClusters);
    const std::uint32_t frontierCap = FrontierCapacity(n, S);
    const std::size_t inner = n - 1 + static_cast<std::size_t>(frontierCap);  // BVH2 inner nodes incl. SAH top
    const std::uint32_t nodeCap = n - 1;  // BVH4 nodes: at most one per reachable BVH2 inner node
    morton_.Reserve(sizeof(std::uint32_t) * n);
    mortonSorted_.Reserve(sizeof(std::uint32_t) * n);
    ids_.Reserve(sizeof(std::uint32_t) * n);
    idsSorted_.Reserve(sizeof(std::uint32_t) * n);
    triBox_.Reserve(sizeof(Box4) * n);
    child_.Reserve(sizeof(std::uint32_t) * 2 * inner);
    parent_.Reserve(sizeof(std::uint32_t) * n);
    leafParent_.Reserve(sizeof(std::uint32_t) * n);
    range_.Reserve(sizeof(std::uint32_t) * 2 * inner);
    arrivals_.Reserve(sizeof(std::uint32_t) * n);
    nodeBox_.Reserve(sizeof(Box4) * inner);
    nodeCost_.Reserve(sizeof(float) * n);
    leafFlag_.Reserve(sizeof(std::uint32_t) * inner);
    source_.Reserve(sizeof(std::uint32_t) * nodeCap);
    nodeStack_.Reserve(sizeof(std::uint32_t) * nodeCap);
    counters_.Reserve(64);
    if (frontierCap) frontier_.Reserve(sizeof(FrontierItem) * frontierCap);
    nodes.Reserve(sizeof(Bvh4Node) * nodeCap);
    slots.Reserve(sizeof(float) * 12 * static_cast<std::size_t>(n));

    const float* pos = reinterpret_cast<const float*>(positions);
    auto* keys = morton_.as<std::uint32_t>();
    auto* sortedKeys = mortonSorted_.as<std::uint32_t>();
    auto* ids = ids_.as<std::uint32_t>();
    auto* sortedIds = idsSorted_.as<std::uint32_t>();
    auto* triBox = triBox_.as<Box4>();
    auto* child2 = child_.as<std::uint32_t>();
    auto* range2 = range_.as<std::uint32_t>();
    auto* nodeBox = nodeBox_.as<Box4>();
    auto* counters = counters_.as<std::uint32_t>();

    float scale[3];
    lbvh::MortonScale(lo, hi, scale);
    const unsigned g = Grid(n);
    HDCX_LAUNCH(prepareKernel, g, pos, indices, n, lo[0], lo[1], lo[2], scale[0], scale[1], scale[2], keys, ids, triBox);
    SortPairs30(keys, ids, sortedKeys, sortedIds, n, st);
    const std::uint32_t none = lbvh::kNone;
    Check(hipMemcpyAsync(parent_.as<void>(), &none, 4, hipMemcpyHostToDevice, st), "lbvh root");
    HDCX_LAUNCH(radixTreeKernel, g, sortedKeys, n, child2, parent_.as<std::uint32_t>(),
                leafParent_.as<std::uint32_t>(), range2);
    Check(hipMemsetAsync(arrivals_.as<void>(), 0, sizeof(std::uint32_t) * n, st), "lbvh arrivals");
    // Leaf flags of the SAH top nodes (ids >= n - 1) stay 0.
    Check(hipMemsetAsync(leafFlag_.as<void>(), 0, sizeof(std::uint32_t) * inner, st), "lbvh leaf flags");
    HDCX_LAUNCH(bottomUpKernel, g, n, child2, parent_.as<std::uint32_t>(), leafParent_.as<std::uint32_t>(),
                arrivals_.as<std::uint32_t>(), sortedIds, triBox, range2, nodeBox, nodeCost_.as<float>(),
                leafFlag_.as<std::uint32_t>());

    // SAH top levels over the frontier.
    std::uint32_t root = 0;
    if (S && n > S) {
        // A frontier that outgrows its capacity under the area bound (many
        // large triangles) is taken again without it.
        std::uint32_t count = 0;
        for (const float maxArea : {FrontierMaxArea(lo, hi, n, S), INFINITY}) {
            Check(hipMemsetAsync(counters, 0, 64, st), "lbvh counters");
            HDCX_LAUNCH(frontierKernel, g, n, S, maxArea, child2, range2, nodeBox, triBox, sortedIds, counters,
                        frontierCap, frontier_.as<FrontierItem>());
            Check(hipMemcpyAsync(&count, counters, 4, hipMemcpyDeviceToHost, st), "lbvh frontier");
            Check(hipStreamSynchronize(st), "lbvh frontier");
            if (count <= frontierCap) break;
        }
        if (count >= 2 && count <= frontierCap) {
            std::vector<FrontierItem> items(count);
            Check(hipMemcpy(items.data(), frontier_.as<void>(), sizeof(FrontierItem) * count, hipMemcpyDeviceToHost),
                  "lbvh frontier");
            // Appended in arrival order: sort for a deterministic top.
            std::sort(items.begin(), items.end(),
                      [](const FrontierItem& a, const FrontierItem& b) { return a.ref < b.ref; });
            std::vector<std::uint32_t> tc, tr;
            std::vector<Box4> tb;
            root = BuildSahTop(items, n - 1, tc, tr, tb);
            const std::size_t base = n - 1;
            Check(hipMemcpyAsync(child2 + 2 * base, tc.data(), sizeof(std::uint32_t) * tc.size(), hipMemcpyHostToDevice,
                                 st),
                  "lbvh top");
            Check(hipMemcpyAsync(range2 + 2 * base, tr.data(), sizeof(std::uint32_t) * tr.size(), hipMemcpyHostToDevice,
                                 st),
                  "lbvh top");
            Check(hipMemcpyAsync(nodeBox + base, tb.data(), sizeof(Box4) * tb.size(), hipMemcpyHostToDevice, st),
                  "lbvh top");
            Check(hipStreamSynchronize(st), "lbvh top");  // tc / tr / tb leave scope
            stats.topClusters = count;
        }
    }

    // Collapse, one level per launch. counters: [0] nodes allocated,
    // [1] max stack, [2] depth, [3] overflow flag.
    const std::uint32_t init[4] = {1U, 0U, 0U, 0U};
    const std::uint32_t zero = 0U;
    Check(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, st), "lbvh counters");
    Check(hipMemcpyAsync(source_.as<void>(), &root, 4, hipMemcpyHostToDevice, st), "lbvh root");
    Check(hipMemcpyAsync(nodeStack_.as<void>(), &zero, 4, hipMemcpyHostToDevice, st), "lbvh root");
    plan.order.Reset();
    plan.begin.assign(1, 0U);
    std::uint32_t begin = 0, end = 1, level = 0;
    std::uint32_t c[4] = {0, 0, 0, 0};
    while (begin < end) {
        HDCX_LAUNCH(collapseKernel, Grid(end - begin), begin, end, level, source_.as<std::uint32_t>(), child2, range2,
                    nodeBox, triBox, sortedIds, leafFlag_.as<std::uint32_t>(), nodeStack_.as<std::uint32_t>(),
                    counters, nodeCap, nodes.as<Bvh4Node>());
        Check(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, st), "lbvh level");
        Check(hipStreamSynchronize(st), "lbvh level");
        if (c[3] != 0U || c[0] > nodeCap) throw Failure(HDCX_ERR_INTERNAL, "device BVH collapse overflowed its node pool");
        plan.begin.push_back(end);
        begin = end;
        end = c[0];
        ++level;
    }
    HDCX_LAUNCH(slotsKernel, g, n, sortedIds, pos, indices, slots.as<float>());
    Check(hipStreamSynchronize(st), "lbvh build");
    stats.nodeCount = c[0];
    stats.maxStack = c[1];
    stats.depth = c[2];
}

void RefitBvh4(const float4* positions, const std::uint32_t* indices, std::uint32_t slotCount, Bvh4Node* nodes,
               float4* slots, const RefitPlan& plan, hipStream_t st)
{
    const float* pos = reinterpret_cast<const float*>(positions);
    auto* s = reinterpret_cast<float*>(slots);
    HDCX_LAUNCH(refitSlotsKernel, Grid(slotCount), slotCount, pos, indices, s);
    const std::uint32_t* order = plan.order.bytes() ? plan.order.as<std::uint32_t>() : nullptr;
    for (std::size_t l = plan.begin.size() - 1; l-- > 0;) {
        const std::uint32_t b = plan.begin[l], e = plan.begin[l + 1];
        if (e > b) HDCX_LAUNCH(refitLevelKernel, Grid(e - b), order, b, e, nodes, static_cast<const float*>(s));
    }
    Check(hipStreamSynchronize(st), "refit");
}

void MakeRefitPlan(const std::vector<Bvh4Node>& nodes, hipStream_t st, RefitPlan& plan)
{
    plan.begin.assign(1, 0U);
    plan.order.Reset();
    if (nodes.empty()) return;
    std::vector<std::uint32_t> order{0U};
    std::size_t b = 0;
    while (b < order.size()) {
        const std::size_t e = order.size();
        plan.begin.push_back(static_cast<std::uint32_t>(e));
        for (std::size_t i = b; i < e; ++i)
            for (std::uint32_t c : nodes[order[i]].child)
                if (c != kEmptyChild && (c & kLeafFlag) == 0) order.push_back(c);
        b = e;
    }
    plan.order.Upload(order.data(), order.size() * sizeof(std::uint32_t), st);
    Check(hipStreamSynchronize(st), "refit plan");
}

} // namespace hdcx
