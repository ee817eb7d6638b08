// Host parts of the device BVH builder: the binned-SAH top levels built
// over the frontier subtrees, and a serial CPU run of every device step
// (same per-element code, lbvh_core.h) used to check the builder without a
// GPU.
#pragma once

#include "bvh.hpp"
#include "lbvh_core.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace hdcx {

// SAH binary tree over the frontier items: at every node the best split
// among all positions of the items sorted by box centre along each axis
// (full sweep; cost = area x triangle count on each side). Appends its
// inner nodes as BVH2 node ids firstId, firstId + 1, ... in depth-first
// pre-order (child2 / range2: two words per node, box: one Box4 per node)
// and returns the root reference. count >= 2.
std::uint32_t BuildSahTop(const std::vector<lbvh::FrontierItem>& items, std::uint32_t firstId,
                          std::vector<std::uint32_t>& child2, std::vector<std::uint32_t>& range2,
                          std::vector<lbvh::Box4>& box);

struct LbvhBuildInfo {
    std::uint32_t topClusters = 0;
    std::vector<std::uint32_t> levelBegin;  // BVH4 ids of level l: [levelBegin[l], levelBegin[l + 1])
};

// The device builder's steps run one element at a time on the host.
// positions: float4 per vertex; indices: 3 per triangle; n >= 2.
Bvh4 BuildBvh4LbvhSerial(const float* positions, const std::uint32_t* indices, std::uint32_t n,
                         std::uint32_t targetClusters, LbvhBuildInfo* info = nullptr);

// BVH4 over boxes (binned SAH, one item per leaf: leaf k = item k), for
// the instance tree. maxStack / depth filled in. items >= 1.
Bvh4 BuildBvh4OverBoxes(const std::vector<lbvh::Box4>& boxes);

// Morton quantisation bounds of a vertex set (the same on both paths).
void VertexBounds(const float* positions, std::uint32_t vertexCount, float lo[3], float hi[3]);

// Frontier size threshold and capacity for n triangles and a target count.
// A scene of no more triangles than the target gets S = 1: the SAH top
// then covers every triangle (a Morton cluster that holds one large
// triangle, a ground quad beside small objects, would give all of its
// nodes the large box).
inline std::uint32_t FrontierThreshold(std::uint32_t n, std::uint32_t targetClusters)
{
    if (targetClusters == 0) return 0;
    return (n + targetClusters - 1) / targetClusters;
}
inline std::uint32_t FrontierCapacity(std::uint32_t n, std::uint32_t S) { return S ? 4U * (n / S) + 8U : 0U; }

// Largest half-area of a frontier item of S of the n triangles in the box
// lo..hi: sqrt(S / n) of the box's half-area. Clusters on a surface cover
// about S / n of its area, so only a cluster that holds a triangle far
// larger than its neighbours is split further.
inline float FrontierMaxArea(const float lo[3], const float hi[3], std::uint32_t n, std::uint32_t S)
{
    const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    return static_cast<float>((dx * dy + dy * dz + dz * dx) * std::sqrt(double(S) / double(n)));
}

} // namespace hdcx
