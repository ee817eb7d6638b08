// Per-element steps of the device BVH builder, written once for the GPU
// kernels (lbvh.hip) and for a serial CPU run of the same code (the
// builder's emulation test). Plain arrays in, plain arrays out; the only
// GPU-specific parts (atomics) come in through a policy type.
#pragma once

#include "bvh.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define HDCX_HD __host__ __device__
#else
#define HDCX_HD
#endif

// No FMA contraction: the device and the serial host run must agree bit for
// bit (Morton codes, and the area comparisons that pick children to open).
#if defined(__clang__)
#define HDCX_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define HDCX_NO_CONTRACT
#endif

namespace hdcx {
namespace lbvh {

// BVH2 child reference: an inner node id, or sorted position | kLeafRef.
constexpr std::uint32_t kLeafRef = 0x80000000U;
constexpr std::uint32_t kNone = 0xffffffffU;
constexpr std::uint32_t kLeafTriangles = 4;       // largest subtree that may become a BVH4 leaf
// Leaf decision (surface-area heuristic, areas relative to the parent's):
// cost of a node visit and of one triangle test.
constexpr float kCostNode = 1.0f, kCostTriangle = 1.0f;
constexpr std::uint32_t kTopRangeHi = 0xfffffffeU;  // range.y of an SAH top node: never a leaf

struct alignas(16) Box4 {
    float lo[4];
    float hi[4];
};

struct FrontierItem {
    std::uint32_t ref, size, pad0, pad1;
    Box4 box;
};

HDCX_HD inline float BitsToFloat(std::uint32_t b)
{
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}
HDCX_HD inline std::uint32_t FloatToBits(float f)
{
    std::uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    return b;
}

HDCX_HD inline std::uint32_t ExpandBits(std::uint32_t v)
{
    v = (v * 0x00010001U) & 0xFF0000FFU;
    v = (v * 0x00000101U) & 0x0F00F00FU;
    v = (v * 0x00000011U) & 0xC30C30C3U;
    v = (v * 0x00000005U) & 0x49249249U;
    return v;
}

// Two ulps toward -inf, so that boxes stay conservative after rounding.
HDCX_HD inline float PadDown(float x)
{
    for (int i = 0; i < 2; ++i) {
        if (x == 0.0f) {
            x = -BitsToFloat(1U);
        } else {
            const std::uint32_t b = FloatToBits(x);
            x = BitsToFloat(x > 0.0f ? b - 1U : b + 1U);
        }
    }
    return x;
}
HDCX_HD inline float PadUp(float x) { return -PadDown(-x); }

HDCX_HD inline float Min3(float a, float b, float c) { return fminf(fminf(a, b), c); }
HDCX_HD inline float Max3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

// Quantisation of box centres for the Morton codes: one scale for all three
// axes (1024 cells over the longest extent of the scene box), so that the
// grid cells are cubes and a flat or elongated scene is split along its
// long axes first. A scale per axis would give a thin axis as many code
// bits as a long one: on a 1024^2 heightfield that raises the tree's SAH
// cost by a fifth.
inline void MortonScale(const float lo[3], const float hi[3], float scale[3])
{
    float extent = 0.0f;
    for (int k = 0; k < 3; ++k) extent = fmaxf(extent, hi[k] - lo[k]);
    for (int k = 0; k < 3; ++k) scale[k] = extent > 0.0f ? 1024.0f / extent : 0.0f;
}

// Box and 30-bit Morton code of triangle i (scale from MortonScale).
HDCX_HD inline std::uint32_t PrepareTriangle(const float* pos, const std::uint32_t* idx, std::uint32_t i,
                                             const float lo[3], const float scale[3], Box4& box)
{
    HDCX_NO_CONTRACT
    const float* a = pos + 4 * static_cast<std::size_t>(idx[3 * static_cast<std::size_t>(i)]);
    const float* b = pos + 4 * static_cast<std::size_t>(idx[3 * static_cast<std::size_t>(i) + 1]);
    const float* c = pos + 4 * static_cast<std::size_t>(idx[3 * static_cast<std::size_t>(i) + 2]);
    std::uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        box.lo[k] = Min3(a[k], b[k], c[k]);
        box.hi[k] = Max3(a[k], b[k], c[k]);
        const float centre = 0.5f * (box.lo[k] + box.hi[k]);
        q[k] = static_cast<std::uint32_t>(fminf(fmaxf((centre - lo[k]) * scale[k], 0.0f), 1023.0f));
    }
    box.lo[3] = box.hi[3] = 0.0f;
    return (ExpandBits(q[0]) << 2) | (ExpandBits(q[1]) << 1) | ExpandBits(q[2]);
}

HDCX_HD inline int Clz32(std::uint32_t x) { return x == 0U ? 32 : __builtin_clz(x); }

// Length of the common prefix of sorted keys i and j; -1 outside [0, n).
// Equal codes are told apart by position (Karras 2012, section 4).
HDCX_HD inline int Delta(const std::uint32_t* keys, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const std::uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + Clz32(static_cast<std::uint32_t>(i ^ j));
    return Clz32(a ^ b);
}

// Inner node i (0 <= i < n - 1) of the binary radix tree over n >= 2
// sorted keys: its children and covered range [lo, hi], and the parent
// links of both children (Karras 2012, figure 4).
HDCX_HD inline void RadixTreeNode(const std::uint32_t* keys, int n, int i, std::uint32_t* child2,
                                  std::uint32_t* parent, std::uint32_t* leafParent, std::uint32_t* range2)
{
    const int d = Delta(keys, n, i, i + 1) - Delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = Delta(keys, n, i, i - d);
    int lmax = 2;
    while (Delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (Delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = Delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (Delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const std::uint32_t left = lo == gamma ? (static_cast<std::uint32_t>(gamma) | kLeafRef)
                                           : static_cast<std::uint32_t>(gamma);
    const std::uint32_t right = hi == gamma + 1 ? (static_cast<std::uint32_t>(gamma + 1) | kLeafRef)
                                                : static_cast<std::uint32_t>(gamma + 1);
    child2[2 * i] = left;
    child2[2 * i + 1] = right;
    range2[2 * i] = static_cast<std::uint32_t>(lo);
    range2[2 * i + 1] = static_cast<std::uint32_t>(hi);
    if (left & kLeafRef) leafParent[left & ~kLeafRef] = static_cast<std::uint32_t>(i);
    else parent[left] = static_cast<std::uint32_t>(i);
    if (right & kLeafRef) leafParent[right & ~kLeafRef] = static_cast<std::uint32_t>(i);
    else parent[right] = static_cast<std::uint32_t>(i);
}

HDCX_HD inline void RefBox(std::uint32_t ref, const Box4* nodeBox, const Box4* triBox, const std::uint32_t* ids,
                           Box4& out)
{
    out = (ref & kLeafRef) ? triBox[ids[ref & ~kLeafRef]] : nodeBox[ref];
}

HDCX_HD inline std::uint32_t RefSize(std::uint32_t ref, const std::uint32_t* range2)
{
    if (ref & kLeafRef) return 1U;
    return range2[2 * ref + 1] - range2[2 * ref] + 1U;
}

HDCX_HD inline float HalfArea(const Box4& b)
{
    HDCX_NO_CONTRACT
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Walks up from leaf k; at each parent the first arrival stops and the
// second writes the union of both children and the subtree's SAH cost:
// a subtree of <= kLeafTriangles triangles is flagged as a leaf when
// testing all of its triangles costs no more than one more node visit
// plus its children's costs (the bottom-up leaf choice of Karras and
// Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies", HPG 2013, section 4). Atomics::Arrive(counter) returns the
// previous count with acquire/release ordering.
template <class Atomics>
HDCX_HD inline void BottomUpLeaf(std::uint32_t k, const std::uint32_t* child2, const std::uint32_t* parent,
                                 const std::uint32_t* leafParent, std::uint32_t* arrivals, const std::uint32_t* ids,
                                 const Box4* triBox, const std::uint32_t* range2, Box4* nodeBox, float* nodeCost,
                                 std::uint32_t* leafFlag)
{
    HDCX_NO_CONTRACT
    std::uint32_t p = leafParent[k];
    while (p != kNone) {
        if (Atomics::Arrive(&arrivals[p]) == 0U) return;
        Box4 a, b;
        RefBox(child2[2 * p], nodeBox, triBox, ids, a);
        RefBox(child2[2 * p + 1], nodeBox, triBox, ids, b);
        Box4 u;
        for (int x = 0; x < 4; ++x) {
            u.lo[x] = fminf(a.lo[x], b.lo[x]);
            u.hi[x] = fmaxf(a.hi[x], b.hi[x]);
        }
        nodeBox[p] = u;
        const std::uint32_t l = child2[2 * p], r = child2[2 * p + 1];
        const float costL = (l & kLeafRef) ? kCostTriangle * HalfArea(a) : nodeCost[l];
        const float costR = (r & kLeafRef) ? kCostTriangle * HalfArea(b) : nodeCost[r];
        const float area = HalfArea(u);
        const float split = kCostNode * area + costL + costR;
        const std::uint32_t size = range2[2 * p + 1] - range2[2 * p] + 1U;
        const float whole = kCostTriangle * area * static_cast<float>(size);
        const bool leaf = size <= kLeafTriangles && whole <= split;
        nodeCost[p] = leaf ? whole : split;
        leafFlag[p] = leaf ? 1U : 0U;
        p = parent[p];
    }
}

// A subtree the SAH top levels take as one item: a single triangle, or at
// most S triangles in a box of half-area <= maxArea. The area bound keeps
// a large triangle (a ground quad) out of a cluster of small ones, whose
// nodes would all carry its box.
HDCX_HD inline bool FrontierItemRef(std::uint32_t ref, std::uint32_t S, float maxArea, const std::uint32_t* range2,
                                    const Box4* nodeBox)
{
    if (ref & kLeafRef) return true;
    return RefSize(ref, range2) <= S && HalfArea(nodeBox[ref]) <= maxArea;
}

// Frontier elements among the two children of inner node i: items
// (FrontierItemRef) under a parent that is not one. Atomics::Append(count)
// returns a slot; slots >= capacity are dropped (the host sees the count).
template <class Atomics>
HDCX_HD inline void FrontierNode(std::uint32_t i, std::uint32_t S, float maxArea, const std::uint32_t* child2,
                                 const std::uint32_t* range2, const Box4* nodeBox, const Box4* triBox,
                                 const std::uint32_t* ids, std::uint32_t* count, std::uint32_t capacity,
                                 FrontierItem* out)
{
    if (FrontierItemRef(i, S, maxArea, range2, nodeBox)) return;
    for (int k = 0; k < 2; ++k) {
        const std::uint32_t ref = child2[2 * i + k];
        const std::uint32_t size = RefSize(ref, range2);
        if (!FrontierItemRef(ref, S, maxArea, range2, nodeBox)) continue;
        const std::uint32_t slot = Atomics::Append(count);
        if (slot >= capacity) continue;
        FrontierItem f;
        f.ref = ref;
        f.size = size;
        f.pad0 = f.pad1 = 0U;
        RefBox(ref, nodeBox, triBox, ids, f.box);
        out[slot] = f;
    }
}

// BVH4 node `id` from BVH2 node source[id]: opens the largest-area inner
// child that is not flagged as a leaf (leafFlag, null for none) until four
// children; flagged children become leaves over their triangle range, the
// other inner children get new node ids from Atomics::Append(&counters[0])
// (counters[3] flags an overflow of `capacity`). counters[1] / [2]: max
// stack, depth.
template <class Atomics>
HDCX_HD inline void CollapseNode(std::uint32_t id, std::uint32_t level, std::uint32_t* source, const std::uint32_t* child2,
                                 const std::uint32_t* range2, const Box4* nodeBox, const Box4* triBox,
                                 const std::uint32_t* ids, const std::uint32_t* leafFlag, std::uint32_t* nodeStack,
                                 std::uint32_t* counters, std::uint32_t capacity, Bvh4Node* nodes)
{
    const std::uint32_t src = source[id];
    std::uint32_t kids[4];
    kids[0] = child2[2 * src];
    kids[1] = child2[2 * src + 1];
    int k = 2;
    while (k < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int i = 0; i < k; ++i) {
            if ((kids[i] & kLeafRef) || (leafFlag && leafFlag[kids[i]])) continue;
            const float a = HalfArea(nodeBox[kids[i]]);
            if (a > bestArea) {
                bestArea = a;
                best = i;
            }
        }
        if (best < 0) break;
        const std::uint32_t open = kids[best];
        kids[best] = child2[2 * open];
        kids[k++] = child2[2 * open + 1];
    }
    Bvh4Node node;
    for (int i = 0; i < 4; ++i) {
        node.lox[i] = node.loy[i] = node.loz[i] = INFINITY;
        node.hix[i] = node.hiy[i] = node.hiz[i] = -INFINITY;
        node.child[i] = kEmptyChild;
        node.pad[i] = 0U;
    }
    const std::uint32_t stack = nodeStack[id] + static_cast<std::uint32_t>(k - 1);
    for (int i = 0; i < k; ++i) {
        Box4 b;
        RefBox(kids[i], nodeBox, triBox, ids, b);
        node.lox[i] = PadDown(b.lo[0]);
        node.loy[i] = PadDown(b.lo[1]);
        node.loz[i] = PadDown(b.lo[2]);
        node.hix[i] = PadUp(b.hi[0]);
        node.hiy[i] = PadUp(b.hi[1]);
        node.hiz[i] = PadUp(b.hi[2]);
        const std::uint32_t size = RefSize(kids[i], range2);
        if (kids[i] & kLeafRef) {
            node.child[i] = EncodeLeaf(kids[i] & ~kLeafRef, 1U);
        } else if (leafFlag && leafFlag[kids[i]]) {
            node.child[i] = EncodeLeaf(range2[2 * kids[i]], size);
        } else {
            const std::uint32_t nid = Atomics::Append(&counters[0]);
            if (nid < capacity) {
                node.child[i] = nid;
                source[nid] = kids[i];
                nodeStack[nid] = stack;
            } else {
                Atomics::Flag(&counters[3]);
            }
        }
    }
    nodes[id] = node;
    Atomics::Max(&counters[1], stack);
    Atomics::Max(&counters[2], level + 1U);
}

// Node boxes of a BVH4 node from its leaves' slots or its child nodes'
// (already refitted) boxes.
HDCX_HD inline void RefitNode(Bvh4Node& node, const Bvh4Node* nodes, const float* slots)
{
    for (int i = 0; i < 4; ++i) {
        const std::uint32_t c = node.child[i];
        if (c == kEmptyChild) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (c & kLeafFlag) {
            const std::uint32_t first = (c >> 3) & 0x0fffffffU, cnt = (c & 7U) + 1U;
            for (std::uint32_t j = 0; j < cnt; ++j)
                for (int v = 0; v < 3; ++v) {
                    const float* p = slots + 12 * static_cast<std::size_t>(first + j) + 4 * v;
                    for (int x = 0; x < 3; ++x) {
                        lo[x] = fminf(lo[x], p[x]);
                        hi[x] = fmaxf(hi[x], p[x]);
                    }
                }
        } else {
            // Child boxes were padded when written; their union is padded
            // again (conservative, a few ulps looser than a rebuild).
            const Bvh4Node& ch = nodes[c];
            for (int j = 0; j < 4; ++j) {
                if (ch.child[j] == kEmptyChild) continue;
                lo[0] = fminf(lo[0], ch.lox[j]);
                lo[1] = fminf(lo[1], ch.loy[j]);
                lo[2] = fminf(lo[2], ch.loz[j]);
                hi[0] = fmaxf(hi[0], ch.hix[j]);
                hi[1] = fmaxf(hi[1], ch.hiy[j]);
                hi[2] = fmaxf(hi[2], ch.hiz[j]);
            }
        }
        node.lox[i] = PadDown(lo[0]);
        node.loy[i] = PadDown(lo[1]);
        node.loz[i] = PadDown(lo[2]);
        node.hix[i] = PadUp(hi[0]);
        node.hiy[i] = PadUp(hi[1]);
        node.hiz[i] = PadUp(hi[2]);
    }
}

// Slot k = triangle t: its three vertices with t in v0.w.
HDCX_HD inline void WriteSlot(const float* pos, const std::uint32_t* idx, std::uint32_t t, float* slot)
{
    for (int v = 0; v < 3; ++v) {
        const float* p = pos + 4 * static_cast<std::size_t>(idx[3 * static_cast<std::size_t>(t) + v]);
        slot[4 * v] = p[0];
        slot[4 * v + 1] = p[1];
        slot[4 * v + 2] = p[2];
        slot[4 * v + 3] = 0.0f;
    }
    slot[3] = BitsToFloat(t);
}

} // namespace lbvh
} // namespace hdcx
