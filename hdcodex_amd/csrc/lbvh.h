// Device BVH construction and refit (gfx950). Produces the same four-wide
// node and triangle-slot layout as the host builder (bvh.hpp), so traversal
// does not know which builder ran.
#pragma once

#include "bvh.hpp"
#include "device_buffer.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hdcx {

// Node ids in bottom-up refit order: level l (root = level 0) is
// order[begin[l] .. begin[l + 1]); an empty `order` means the identity
// (the device builder allocates each level as one contiguous id range).
struct RefitPlan {
    DeviceBuffer order;
    std::vector<std::uint32_t> begin;
};

struct DeviceBvhStats {
    std::uint32_t nodeCount = 0;
    std::uint32_t maxStack = 0;
    std::uint32_t depth = 0;
    std::uint32_t topClusters = 0;  // subtrees under the SAH-built top levels
};

class DeviceBvhBuilder {
public:
    // Builds the BVH4 of n >= 2 triangles. positions: float4 per vertex,
    // indices: 3 per triangle, both on the device; lo / hi bound every
    // vertex. Writes the nodes (root at 0) and n triangle slots into the
    // given buffers and the refit plan; synchronises `st`.
    // targetClusters: number of Morton subtrees under the binned-SAH top
    // levels (0 = pure LBVH).
    void Build(const float4* positions, const std::uint32_t* indices, std::uint32_t n, const float lo[3],
               const float hi[3], std::uint32_t targetClusters, hipStream_t st, DeviceBuffer& nodes,
               DeviceBuffer& slots, RefitPlan& plan, DeviceBvhStats& stats);

private:
    // Stable radix sort of (Morton code, triangle id) pairs; the result is in
    // (sa, sb), a and b are overwritten.
    void SortPairs30(std::uint32_t* a, std::uint32_t* b, std::uint32_t* sa, std::uint32_t* sb, std::uint32_t n,
                     hipStream_t st);
    DeviceBuffer morton_, mortonSorted_, ids_, idsSorted_, triBox_, child_, parent_, leafParent_, range_, arrivals_,
        nodeBox_, nodeCost_, leafFlag_, source_, nodeStack_, counters_, frontier_, frontierBox_, temp_;
};

// Re-reads the slot vertices from `positions` (same triangles, moved
// vertices) and recomputes every node box bottom-up; synchronises `st`.
void RefitBvh4(const float4* positions, const std::uint32_t* indices, std::uint32_t slotCount, Bvh4Node* nodes,
               float4* slots, const RefitPlan& plan, hipStream_t st);

// Level-order plan of a BVH4 held on the host (any builder).
void MakeRefitPlan(const std::vector<Bvh4Node>& nodes, hipStream_t st, RefitPlan& plan);

} // namespace hdcx
