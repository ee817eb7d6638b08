// Wide-BVH layout shared by the builders and the traversal kernels.
//
// One node is 128 bytes (one L2 cache line on MI355X): the six slab planes
// of its four children as SoA float4 (so one 16-byte load serves four
// children), four child references and 16 bytes of padding. A child
// reference is an inner-node index, or a leaf: bit 31 set, bits 3..30 the
// first triangle slot, bits 0..2 the triangle count minus one (1..8). Empty
// child slots carry an inverted box (lo = +inf, hi = -inf) so that the slab
// test rejects them without a branch.
//
// Leaf triangles are re-laid out in leaf order as three float4 per slot:
// the vertex positions, with the flattened primitive id in v0.w (as bits).
#pragma once

#include <cstdint>
#include <vector>

namespace hdcx {

struct alignas(16) Bvh4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    std::uint32_t child[4];
    std::uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "BVH4 node is one 128-byte line");

constexpr std::uint32_t kLeafFlag = 0x80000000U;
constexpr std::uint32_t kEmptyChild = 0xffffffffU;
constexpr std::uint32_t kMaxLeafTriangles = 8U;

constexpr std::uint32_t EncodeLeaf(std::uint32_t first, std::uint32_t count)
{
    return kLeafFlag | (first << 3) | (count - 1U);
}

struct Bvh4 {
    std::vector<Bvh4Node> nodes;         // node 0 is the root
    std::vector<float> triangles;        // 12 floats per slot (3 x float4)
    std::uint32_t maxStack{0};           // worst-case traversal stack entries
    std::uint32_t depth{0};
};

// Binned-SAH BVH2 over triangles (16 bins, leaves of <= 4 triangles),
// collapsed to a 4-wide BVH. positions: float4 per vertex; indices: 3 per
// triangle. Deterministic for a given input.
Bvh4 BuildBvh4Host(const float* positions, const std::uint32_t* indices, std::uint32_t triangleCount);

// Worst-case stack occupancy and depth of a BVH4 (root at node 0).
void MeasureBvh4(Bvh4& bvh);

// Structural check of a BVH4 against the triangle count: every inner
// reference in range, every leaf inside the slot array with 1..8 triangles,
// every slot's primitive id < triangleCount, every triangle referenced
// exactly once, empty slots only with inverted boxes. Throws
// std::runtime_error with the first violation.
void ValidateBvh4(const Bvh4& bvh, std::uint32_t triangleCount);

} // namespace hdcx
