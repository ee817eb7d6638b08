// Wide-BVH traversal for gfx950: one ray per lane, a four-child node per
// 128-byte fetch, near-to-far ordered stack whose first 16 entries live in
// LDS (lane-interleaved, conflict-free) and the rest in private memory.
//
// Hit semantics (what the reference leaves to the Vulkan intersector,
// restated as a definition): a triangle is hit when the watertight test of
// Woop, Benthin and Wald (JCGT 2013) passes with tmin <= t <= tmax; the
// closest hit is the smallest (t, primitive id) pair. The triangle test runs
// with FP contraction off and IEEE division, so t, u, v are bit-identical to
// any IEEE FP32 host evaluating the same operations.
#pragma once

#include "dev_common.h"

namespace hdcx {

constexpr int kLdsStack = 16;
constexpr int kSpillStack = 48;
constexpr int kMaxStack = kLdsStack + kSpillStack;

struct RayPre {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;  // reciprocal direction for the slab test
    int kx, ky, kz;
    float sx, sy, sz;
    // Byte offset of the entry plane within a node's (lo, hi) plane pair of
    // each axis: 0 (lo) when the reciprocal is positive, 16 (hi) otherwise.
    std::uint32_t nearX, nearY, nearZ;
    // Slab distances are fma(plane, i, c) per axis: c = -o * i, lowered for
    // entry planes (cn*) and raised for exit planes (cf*) by 2^-22 |c|,
    // four times the rounding error of c (see testNode).
    float cnx, cny, cnz, cfx, cfy, cfz;
};

// Entry-plane offsets of a reciprocal direction (see RayPre).
__device__ __forceinline__ std::uint32_t nearPlaneOffset(float inv) { return inv < 0.0f ? 16U : 0U; }

struct HitRec {
    float t;
    std::uint32_t prim;
    float u, v;
};

__device__ __forceinline__ float pick(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ float safeRcp(float d)
{
    // Axis-parallel rays: a huge finite reciprocal keeps the slab test
    // free of 0 * inf.
    const float tiny = 1e-30f;
    return 1.0f / (fabsf(d) > tiny ? d : copysignf(tiny, d));
}

__device__ __forceinline__ RayPre prepareRay(float ox, float oy, float oz, float dx, float dy, float dz)
{
    RayPre r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = safeRcp(dx); r.iy = safeRcp(dy); r.iz = safeRcp(dz);
    r.nearX = nearPlaneOffset(r.ix); r.nearY = nearPlaneOffset(r.iy); r.nearZ = nearPlaneOffset(r.iz);
    {
        const float cx = -ox * r.ix, cy = -oy * r.iy, cz = -oz * r.iz;
        const float ex = fabsf(cx) * 0x1p-22f, ey = fabsf(cy) * 0x1p-22f, ez = fabsf(cz) * 0x1p-22f;
        r.cnx = cx - ex; r.cny = cy - ey; r.cnz = cz - ez;
        r.cfx = cx + ex; r.cfy = cy + ey; r.cfz = cz + ez;
    }
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = pick(dx, dy, dz, kz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.sx = pick(dx, dy, dz, kx) / dkz;
    r.sy = pick(dx, dy, dz, ky) / dkz;
    r.sz = 1.0f / dkz;
    return r;
}

// Watertight ray/triangle test; returns true with t in [tmin, tmax]. The
// barycentrics are u = V / det and v = W / det; the closest-hit loops keep
// V, W and det of the best candidate and divide once at the end (the same
// quotients, without two divisions per accepted candidate).
__device__ __forceinline__ bool intersectTriangle(const RayPre& r, const float4& p0, const float4& p1,
                                                  const float4& p2, float tmin, float tmax, float& t, float& vNum,
                                                  float& wNum, float& det)
{
#pragma clang fp contract(off)
    const float ax0 = p0.x - r.ox, ay0 = p0.y - r.oy, az0 = p0.z - r.oz;
    const float bx0 = p1.x - r.ox, by0 = p1.y - r.oy, bz0 = p1.z - r.oz;
    const float cx0 = p2.x - r.ox, cy0 = p2.y - r.oy, cz0 = p2.z - r.oz;
    const float aK = pick(ax0, ay0, az0, r.kz), bK = pick(bx0, by0, bz0, r.kz), cK = pick(cx0, cy0, cz0, r.kz);
    const float ax = pick(ax0, ay0, az0, r.kx) - r.sx * aK, ay = pick(ax0, ay0, az0, r.ky) - r.sy * aK;
    const float bx = pick(bx0, by0, bz0, r.kx) - r.sx * bK, by = pick(bx0, by0, bz0, r.ky) - r.sy * bK;
    const float cx = pick(cx0, cy0, cz0, r.kx) - r.sx * cK, cy = pick(cx0, cy0, cz0, r.ky) - r.sy * cK;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double cxd = cx, cyd = cy, bxd = bx, byd = by, axd = ax, ayd = ay;
        U = static_cast<float>(cxd * byd - cyd * bxd);
        V = static_cast<float>(axd * cyd - ayd * cxd);
        W = static_cast<float>(bxd * ayd - byd * axd);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    det = U + V + W;
    if (det == 0.0f) return false;
    const float az = r.sz * aK, bz = r.sz * bK, cz = r.sz * cK;
    const float T = U * az + V * bz + W * cz;
    t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    vNum = V;
    wNum = W;
    return true;
}

// Slab test of the four children of a node; returns near distances
// (+inf when missed).
//
// The node stores each axis as a lo plane and a hi plane of the four
// children (16 bytes each). The ray's direction signs say which of the two
// it enters through, so the entry and exit planes are loaded directly
// (RayPre::near*) and their distances need no min/max: for a positive
// reciprocal the entry plane is lo, for a negative one hi.
//
// Each plane distance is one fma, fma(p, i, c), instead of (p - o) * i.
// Against the exact (p - o) * i (with the rounded reciprocal) the fma's
// only errors are c's rounding, at most 2^-24 |c|, and its own final
// rounding, 2^-24 of the result. The entry constant is c - 2^-22 |c| and the
// exit constant c + 2^-22 |c| (both still beyond -o * i after their own
// rounding), so entry distances are never above and exit distances never
// below the exact ones except by the final rounding, which the 4-ulp
// inflation of tfar covers. The test therefore never culls a box the exact
// test keeps; which boxes it keeps only changes the order and number of
// triangle tests, and the closest hit (ties broken by primitive id) does
// not depend on either. A NaN distance (an overflowing constant) drops out
// of the min/max, which keeps the box.
//
// Node addresses are a 32-bit byte offset (index x 128) from the array base
// (node arrays stay below 4 GiB: the host caps them), which the compiler
// keeps in one register per load next to the scalar base.
__device__ __forceinline__ const float4& nodePlanes(const Bvh4Node* __restrict__ nodes, std::uint32_t byteOffset)
{
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(nodes) + byteOffset);
}
__device__ __forceinline__ void testNode(const Bvh4Node* __restrict__ nodes, std::uint32_t index, const RayPre& r, float tmin,
                                         float tmax, float tn[4], std::uint32_t ref[4])
{
    const std::uint32_t base = index << 7;
    const float4 nx = nodePlanes(nodes, base + r.nearX);
    const float4 fx = nodePlanes(nodes, base + (16U - r.nearX));
    const float4 ny = nodePlanes(nodes, base + 32U + r.nearY);
    const float4 fy = nodePlanes(nodes, base + (48U - r.nearY));
    const float4 nz = nodePlanes(nodes, base + 64U + r.nearZ);
    const float4 fz = nodePlanes(nodes, base + (80U - r.nearZ));
    const uint4 ch = *reinterpret_cast<const uint4*>(&nodePlanes(nodes, base + 96U));
    const float nxs[4] = {nx.x, nx.y, nx.z, nx.w}, fxs[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nys[4] = {ny.x, ny.y, ny.z, ny.w}, fys[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nzs[4] = {nz.x, nz.y, nz.z, nz.w}, fzs[4] = {fz.x, fz.y, fz.z, fz.w};
    const std::uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float tx0 = fmaf(nxs[i], r.ix, r.cnx), tx1 = fmaf(fxs[i], r.ix, r.cfx);
        const float ty0 = fmaf(nys[i], r.iy, r.cny), ty1 = fmaf(fys[i], r.iy, r.cfy);
        const float tz0 = fmaf(nzs[i], r.iz, r.cnz), tz1 = fmaf(fzs[i], r.iz, r.cfz);
        const float tnear = fmaxf(fmaxf(tx0, ty0), fmaxf(tz0, tmin));
        const float tfar = fminf(fminf(tx1, ty1), fminf(tz1, tmax));
        // Empty slots are rejected by reference: their planes are not a box.
        const bool hit = c[i] != kEmptyChild && tnear <= tfar * 1.00000048f;
        tn[i] = hit ? tnear : INFINITY;
        ref[i] = c[i];
    }
}

// The spill entries are a private array of their own, declared by the
// kernel next to the stack (HDCX_TRAVERSAL_STACK): an array indexed at run
// time inside this struct would keep the whole struct, the stack pointer
// included, in scratch memory, with a scratch load and store on every push
// and pop.
struct TraversalStack {
    std::uint32_t* lds;      // this lane's column; stride = block size
    std::uint32_t stride;
    std::uint32_t* spill;    // kSpillStack private entries
    int sp;
    // The host rejects BVHs whose worst-case occupancy exceeds kMaxStack;
    // the bound check only keeps a corrupt tree from writing out of range.
    __device__ __forceinline__ void push(std::uint32_t v)
    {
        if (sp < kLdsStack) lds[sp * stride] = v;
        else if (sp < kMaxStack) spill[sp - kLdsStack] = v;
        else return;
        ++sp;
    }
    __device__ __forceinline__ std::uint32_t pop()
    {
        --sp;
        return sp < kLdsStack ? lds[sp * stride] : spill[sp - kLdsStack];
    }
    // Pushes the valid ones of a node's far children r3, r2, r1 (a sorted
    // node: validity is a prefix, v3 implies v2 implies v1). With room for
    // all three in LDS every entry is written at the top and the top
    // advances past the valid ones, without a branch per child; the write
    // of an invalid child lands on the free slot above the top (at most
    // index sp + 2 <= kLdsStack - 1).
    __device__ __forceinline__ void pushFar3(std::uint32_t r3, bool v3, std::uint32_t r2, bool v2, std::uint32_t r1,
                                             bool v1)
    {
        if (sp + 3 <= kLdsStack) {
            lds[sp * stride] = r3;
            sp += v3 ? 1 : 0;
            lds[sp * stride] = r2;
            sp += v2 ? 1 : 0;
            lds[sp * stride] = r1;
            sp += v1 ? 1 : 0;
        } else {
            if (v3) push(r3);
            if (v2) push(r2);
            if (v1) push(r1);
        }
    }
    // Any-hit order: continues with the first valid child (next, false when
    // none) and pushes the other valid ones in child order, as the loop
    // "first valid -> next, later valid -> push" does. The fast path writes
    // every child at the top and advances past the pushed ones; before the
    // last write at most two were pushed, so it lands at index <= sp + 2.
    __device__ __forceinline__ bool pushAfterFirst(const float tn[4], const std::uint32_t ref[4], std::uint32_t& next)
    {
        bool have = false;
        if (sp + 3 <= kLdsStack) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool valid = tn[i] != INFINITY;
                if (valid && !have) next = ref[i];
                lds[sp * stride] = ref[i];
                sp += (valid && have) ? 1 : 0;
                have = have || valid;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (tn[i] != INFINITY) {
                    if (!have) { next = ref[i]; have = true; }
                    else push(ref[i]);
                }
            }
        }
        return have;
    }
};

// A kernel's traversal stack over its LDS block `mem` (kLdsStack x kBlock
// words): `name` and its private spill array.
#define HDCX_TRAVERSAL_STACK(name, mem)               \
    std::uint32_t name##Spill[kSpillStack];           \
    TraversalStack name;                              \
    name.lds = (mem) + threadIdx.x;                   \
    name.stride = kBlock;                             \
    name.spill = name##Spill;                         \
    name.sp = 0

#define HDCX_SWAP_IF(a, b)                                        \
    if (tn[b] < tn[a]) {                                          \
        const float tt = tn[a]; tn[a] = tn[b]; tn[b] = tt;        \
        const std::uint32_t rr = ref[a]; ref[a] = ref[b]; ref[b] = rr; \
    }

// Near-to-far order of a node's four children (5-comparator network).
__device__ __forceinline__ void sortChildren(float tn[4], std::uint32_t ref[4])
{
    HDCX_SWAP_IF(0, 1) HDCX_SWAP_IF(2, 3) HDCX_SWAP_IF(0, 2) HDCX_SWAP_IF(1, 3) HDCX_SWAP_IF(1, 2)
}

// Closest-hit traversal state of one ray (see closestRound).
struct ClosestState {
    std::uint32_t cur;   // next node or leaf to visit (valid while !exhausted)
    bool exhausted;      // nothing left to visit
    bool found;
    float bestT;
    std::uint32_t bestPrim;
    float bestV, bestW, bestDet;  // barycentric numerators and denominator of the best hit
};

__device__ __forceinline__ void startClosest(ClosestState& c, float tmax, TraversalStack& st)
{
    c.cur = 0;  // root is an inner node
    c.exhausted = false;
    c.found = false;
    c.bestT = tmax;
    c.bestPrim = 0xffffffffU;
    c.bestV = c.bestW = 0.0f;
    c.bestDet = 1.0f;
    st.sp = 0;
}

// Takes the next leaf: `cur` (a leaf) becomes `leaf` and the next stack
// entry becomes `cur` (or the traversal is exhausted).
__device__ __forceinline__ void takeLeaf(std::uint32_t& leaf, ClosestState& c, TraversalStack& st)
{
    leaf = c.cur;
    if (st.sp == 0) c.exhausted = true;
    else c.cur = st.pop();
}

// One round of closest-hit traversal: inner nodes, then the leaves met.
//
// Speculative traversal (Aila and Laine, "Understanding the Efficiency of
// Ray Traversal on GPUs", 2009): the inner-node loop postpones the first leaf
// it meets and keeps descending until every lane still in the loop holds a
// leaf, so node fetches and triangle tests run with more lanes of the wave.
// Nodes visited past a postponed leaf are culled against the closest hit so
// far, which only admits more of them; the closest hit (ties broken by
// primitive id) does not depend on the visiting order. For a lone lane the
// order is that of the plain loop (each leaf as it is met). Only triangles
// with (t, prim) > (loT, loPrim) are accepted (loPrim = 0xffffffff: all).
template <bool kCount>
__device__ __forceinline__ void closestRound(const DeviceScene& s, const RayPre& r, float tmin, float loT,
                                             std::uint32_t loPrim, ClosestState& c, TraversalStack& st,
                                             unsigned long long& nodes, unsigned long long& tris)
{
    std::uint32_t leaf = 0;  // postponed leaf; 0 = none (leaf references carry kLeafFlag)
    while (!c.exhausted && !(c.cur & kLeafFlag)) {
        if (kCount) ++nodes;
        float tn[4];
        std::uint32_t ref[4];
        testNode(s.nodes, c.cur, r, tmin, c.bestT, tn, ref);
        sortChildren(tn, ref);
        if (tn[0] != INFINITY) {
            st.pushFar3(ref[3], tn[3] != INFINITY, ref[2], tn[2] != INFINITY, ref[1], tn[1] != INFINITY);
            c.cur = ref[0];
        } else if (st.sp == 0) {
            c.exhausted = true;
        } else {
            c.cur = st.pop();
        }
        if (!c.exhausted && (c.cur & kLeafFlag) && leaf == 0) takeLeaf(leaf, c, st);
        if (!__any(leaf == 0)) break;
    }
    if (leaf == 0 && !c.exhausted && (c.cur & kLeafFlag)) takeLeaf(leaf, c, st);
    while (leaf != 0) {
        const std::uint32_t first = (leaf >> 3) & 0x0fffffffU;
        const std::uint32_t count = (leaf & 7U) + 1U;
        for (std::uint32_t j = 0; j < count; ++j) {
            const float4* tp = s.tris + static_cast<std::size_t>(first + j) * 3;
            const float4 a = tp[0], b = tp[1], cc = tp[2];
            const std::uint32_t prim = __float_as_uint(a.w);
            if (kCount) ++tris;
            float t, vn, wn, det;
            if (intersectTriangle(r, a, b, cc, tmin, c.bestT, t, vn, wn, det)) {
                const bool aboveLo = loPrim == 0xffffffffU || t > loT || (t == loT && prim > loPrim);
                const bool better = !c.found || t < c.bestT || (t == c.bestT && prim < c.bestPrim);
                if (aboveLo && better) {
                    c.found = true;
                    c.bestT = t;
                    c.bestPrim = prim;
                    c.bestV = vn;
                    c.bestW = wn;
                    c.bestDet = det;
                }
            }
        }
        leaf = 0;
        if (!c.exhausted && (c.cur & kLeafFlag)) takeLeaf(leaf, c, st);
    }
}

// The hit record of a finished traversal: (tmax, none, 0, 0) for a miss.
__device__ __forceinline__ HitRec closestResult(const ClosestState& c)
{
    HitRec h;
    h.t = c.bestT;
    h.prim = c.bestPrim;
    h.u = c.found ? c.bestV / c.bestDet : 0.0f;
    h.v = c.found ? c.bestW / c.bestDet : 0.0f;
    return h;
}

// Closest hit with a lexicographic lower bound: only triangles with
// (t, prim) > (loT, loPrim) are accepted (loPrim = 0xffffffff with
// loT = tmin accepts everything from tmin on). Used directly for opaque
// scenes and iterated for candidate-ordered any-hit in non-opaque ones.
template <bool kCount>
__device__ bool traceClosest(const DeviceScene& s, const RayPre& r, float tmin, float tmax, float loT,
                             std::uint32_t loPrim, HitRec& best, TraversalStack& st, TraversalCounters* ctr)
{
    best.t = tmax;
    best.prim = 0xffffffffU;
    best.u = best.v = 0.0f;
    if (s.nodeCount == 0) return false;
    ClosestState c;
    startClosest(c, tmax, st);
    unsigned long long nodes = 0, tris = 0;
    while (!c.exhausted) closestRound<kCount>(s, r, tmin, loT, loPrim, c, st, nodes, tris);
    best = closestResult(c);
    if (kCount && ctr) {
        atomicAdd(&ctr->nodes, nodes);
        atomicAdd(&ctr->triangles, tris);
    }
    return c.found;
}

// Any hit in [tmin, tmax] (terminate on first hit).
template <bool kCount>
__device__ bool traceAny(const DeviceScene& s, const RayPre& r, float tmin, float tmax, TraversalStack& st,
                         TraversalCounters* ctr)
{
    if (s.nodeCount == 0) return false;
    st.sp = 0;
    std::uint32_t cur = 0;
    bool hitAny = false;
    unsigned long long nodes = 0, tris = 0;
    for (;;) {
        if (cur & kLeafFlag) {
            const std::uint32_t first = (cur >> 3) & 0x0fffffffU;
            const std::uint32_t count = (cur & 7U) + 1U;
            for (std::uint32_t j = 0; j < count; ++j) {
                const float4* tp = s.tris + static_cast<std::size_t>(first + j) * 3;
                if (kCount) ++tris;
                float t, vn, wn, det;
                if (intersectTriangle(r, tp[0], tp[1], tp[2], tmin, tmax, t, vn, wn, det)) {
                    hitAny = true;
                    break;
                }
            }
            if (hitAny) break;
        } else {
            if (kCount) ++nodes;
            float tn[4];
            std::uint32_t ref[4];
            testNode(s.nodes, cur, r, tmin, tmax, tn, ref);
            std::uint32_t next = 0xffffffffU;
            if (st.pushAfterFirst(tn, ref, next)) { cur = next; continue; }
        }
        if (st.sp == 0) break;
        cur = st.pop();
    }
    if (kCount && ctr) {
        atomicAdd(&ctr->nodes, nodes);
        atomicAdd(&ctr->triangles, tris);
    }
    return hitAny;
}

#undef HDCX_SWAP_IF

} // namespace hdcx
