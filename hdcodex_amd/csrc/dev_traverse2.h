// Two-level traversal (instancing). The instance tree (TLAS) is a BVH4 whose
// leaves are items: an instance, or (item == instanceCount) the BVH of the
// drawn triangles. Entering an instance pushes an exit marker and switches
// the box tests to the object-space ray; triangles are still tested in world
// space, on vertices transformed with the same FP32 formula the host uses to
// flatten the scene, so hits equal those of the flattened scene bit for bit.
// The object-space box tests widen every box by a bound on the rounding of
// the ray transform and of the vertex transform, so no box that contains a
// world-space hit is culled.
#pragma once

#include "dev_traverse.h"

namespace hdcx {

constexpr std::uint32_t kInstanceExit = 0xfffffffeU;  // stack marker: back to the instance tree

struct BoxRay {
    float ox, oy, oz;
    float ix, iy, iz;
    float pad;  // widening of every box (object space)
    std::uint32_t nearX, nearY, nearZ;  // entry-plane offsets, as RayPre's
};

__device__ __forceinline__ BoxRay worldBoxRay(const RayPre& r)
{
    return BoxRay{r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, 0.0f, r.nearX, r.nearY, r.nearZ};
}

// Object-space ray of instance i and the box widening for hits up to tmax.
// The rounding of the transformed ray grows with the ray parameter t, and a
// hit on the instance lies inside its world box, so t·|d|∞ is bounded by
// the largest coordinate distance from the origin to that box.
__device__ __forceinline__ BoxRay objectBoxRay(const DeviceScene& s, std::uint32_t i, const RayPre& r, float tmax)
{
    const InstanceRecord& in = s.instances[i];
    float o[3], d[3];
    TransformPoint(in.toObject, r.ox, r.oy, r.oz, o);
    {
#pragma clang fp contract(off)
        for (int k = 0; k < 3; ++k) d[k] = (in.toObject[4 * k] * r.dx + in.toObject[4 * k + 1] * r.dy) + in.toObject[4 * k + 2] * r.dz;
    }
    // invNorm, norm, |T|, object extent; world box lo; world box hi.
    const float4* q = reinterpret_cast<const float4*>(s.instancePad) + 3U * i;
    const float4 p = q[0], lo = q[1], hi = q[2];
    const float oInf = fmaxf(fmaxf(fabsf(r.ox), fabsf(r.oy)), fabsf(r.oz));
    const float dInf = fmaxf(fmaxf(fabsf(r.dx), fabsf(r.dy)), fabsf(r.dz));
    const float reach = fmaxf(fmaxf(fmaxf(fabsf(lo.x - r.ox), fabsf(hi.x - r.ox)), fmaxf(fabsf(lo.y - r.oy), fabsf(hi.y - r.oy))),
                              fmaxf(fabsf(lo.z - r.oz), fabsf(hi.z - r.oz)));
    const float td = fminf(fminf(tmax, 1e30f) * dInf, reach * 1.00001f);
    const float pad = 16.0f * 5.9604645e-8f * (p.x * (oInf + 2.0f * p.z + td + p.y * p.w) + p.w) + 1e-6f * p.w;
    const float ix = safeRcp(d[0]), iy = safeRcp(d[1]), iz = safeRcp(d[2]);
    return BoxRay{o[0], o[1], o[2], ix, iy, iz, pad, nearPlaneOffset(ix), nearPlaneOffset(iy), nearPlaneOffset(iz)};
}

__device__ __forceinline__ void testNodeBox(const Bvh4Node* __restrict__ nodes, std::uint32_t index, const BoxRay& r,
                                            float tmin, float tmax, float tn[4], std::uint32_t ref[4])
{
    // Entry and exit planes loaded by direction sign as in testNode; the
    // widening moves an entry plane against the ray (lo - pad or hi + pad)
    // and an exit plane with it.
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
    const float px = r.nearX ? r.pad : -r.pad, py = r.nearY ? r.pad : -r.pad, pz = r.nearZ ? r.pad : -r.pad;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float tx0 = (nxs[i] + px - r.ox) * r.ix, tx1 = (fxs[i] - px - r.ox) * r.ix;
        const float ty0 = (nys[i] + py - r.oy) * r.iy, ty1 = (fys[i] - py - r.oy) * r.iy;
        const float tz0 = (nzs[i] + pz - r.oz) * r.iz, tz1 = (fzs[i] - pz - r.oz) * r.iz;
        const float tnear = fmaxf(fmaxf(tx0, ty0), fmaxf(tz0, tmin));
        const float tfar = fminf(fminf(tx1, ty1), fminf(tz1, tmax));
        const bool hit = c[i] != kEmptyChild && tnear <= tfar * 1.00000048f;
        tn[i] = hit ? tnear : INFINITY;
        ref[i] = c[i];
    }
}

// World-space vertices of leaf slot k (object space, transformed when inside
// an instance).
__device__ __forceinline__ void slotWorld(const DeviceScene& s, std::uint32_t k, int inst, float4& a, float4& b, float4& c)
{
    const float4* tp = s.tris + static_cast<std::size_t>(k) * 3;
    a = tp[0];
    b = tp[1];
    c = tp[2];
    if (inst >= 0) {
        const float* m = s.instances[inst].toWorld;
        float w[3];
        TransformPoint(m, a.x, a.y, a.z, w);
        a.x = w[0]; a.y = w[1]; a.z = w[2];
        TransformPoint(m, b.x, b.y, b.z, w);
        b.x = w[0]; b.y = w[1]; b.z = w[2];
        TransformPoint(m, c.x, c.y, c.z, w);
        c.x = w[0]; c.y = w[1]; c.z = w[2];
    }
}

#define HDCX_SWAP_IF2(a, b)                                       \
    if (tn[b] < tn[a]) {                                          \
        const float tt = tn[a]; tn[a] = tn[b]; tn[b] = tt;        \
        const std::uint32_t rr = ref[a]; ref[a] = ref[b]; ref[b] = rr; \
    }

// Closest hit over the instance tree, with the same lexicographic lower
// bound as traceClosest. Primitive ids are flattened ids. Speculative as
// traceClosest: a triangle leaf of the current instance's BVH is postponed
// while the lane keeps descending; the postponed leaf is tested before the
// lane leaves the instance (kInstanceExit ends the inner-node loop, and
// instance-tree items are only met after it), so it is always tested in its
// own instance's frame.
template <bool kCount>
__device__ bool traceClosest2L(const DeviceScene& s, const RayPre& r, float tmin, float tmax, float loT,
                               std::uint32_t loPrim, HitRec& best, TraversalStack& st, TraversalCounters* ctr)
{
    best.t = tmax;
    best.prim = 0xffffffffU;
    best.u = best.v = 0.0f;
    if (s.nodeCount == 0) return false;
    bool found = false;
    float bestV = 0.0f, bestW = 0.0f, bestDet = 1.0f;
    st.sp = 0;
    std::uint32_t cur = s.tlasRoot;
    bool exhausted = false;
    bool inBlas = false;
    int inst = -1;
    std::uint32_t base = 0, first = 0;
    BoxRay box = worldBoxRay(r);
    unsigned long long nodes = 0, tris = 0;
    // Next item into `leaf` when it is a triangle leaf of the current BVH.
    auto takeTriangleLeaf = [&](std::uint32_t& leaf) {
        if (!exhausted && inBlas && cur != kInstanceExit && (cur & kLeafFlag)) {
            leaf = cur;
            if (st.sp == 0) exhausted = true;
            else cur = st.pop();
        }
    };
    while (!exhausted) {
        std::uint32_t leaf = 0;  // postponed triangle leaf; 0 = none
        while (!exhausted && !(cur & kLeafFlag)) {
            if (kCount) ++nodes;
            float tn[4];
            std::uint32_t ref[4];
            testNodeBox(s.nodes, cur, box, tmin, best.t, tn, ref);
            HDCX_SWAP_IF2(0, 1) HDCX_SWAP_IF2(2, 3) HDCX_SWAP_IF2(0, 2) HDCX_SWAP_IF2(1, 3) HDCX_SWAP_IF2(1, 2)
            if (tn[0] != INFINITY) {
                st.pushFar3(ref[3], tn[3] != INFINITY, ref[2], tn[2] != INFINITY, ref[1], tn[1] != INFINITY);
                cur = ref[0];
            } else if (st.sp == 0) {
                exhausted = true;
            } else {
                cur = st.pop();
            }
            if (leaf == 0) takeTriangleLeaf(leaf);
            if (!__any(leaf == 0)) break;
        }
        if (leaf == 0) takeTriangleLeaf(leaf);
        while (leaf != 0) {
            const std::uint32_t lf = (leaf >> 3) & 0x0fffffffU;
            const std::uint32_t count = (leaf & 7U) + 1U;
            for (std::uint32_t j = 0; j < count; ++j) {
                float4 a, b, c;
                slotWorld(s, lf + j, inst, a, b, c);
                const std::uint32_t slotPrim = __float_as_uint(s.tris[static_cast<std::size_t>(lf + j) * 3].w);
                const std::uint32_t prim = inst >= 0 ? base + (slotPrim - first) : slotPrim;
                if (kCount) ++tris;
                float t, vn, wn, det;
                if (intersectTriangle(r, a, b, c, tmin, best.t, t, vn, wn, det)) {
                    const bool aboveLo = loPrim == 0xffffffffU || t > loT || (t == loT && prim > loPrim);
                    const bool better = !found || t < best.t || (t == best.t && prim < best.prim);
                    if (aboveLo && better) {
                        found = true;
                        best.t = t;
                        best.prim = prim;
                        bestV = vn;
                        bestW = wn;
                        bestDet = det;
                    }
                }
            }
            leaf = 0;
            takeTriangleLeaf(leaf);
        }
        if (exhausted || !(cur & kLeafFlag)) continue;
        if (cur == kInstanceExit) {
            // Back to the instance tree.
            inBlas = false;
            inst = -1;
            box = worldBoxRay(r);
            if (st.sp == 0) exhausted = true;
            else cur = st.pop();
        } else {
            // Item of the instance tree: enter its BVH.
            const std::uint32_t lf = (cur >> 3) & 0x0fffffffU;
            st.push(kInstanceExit);
            inBlas = true;
            if (lf >= s.instanceCount) {
                inst = -1;
                base = 0;
                first = 0;
                cur = s.drawnRoot;
            } else {
                inst = static_cast<int>(lf);
                const PrototypeRecord pr = s.prototypes[s.instances[lf].prototype];
                base = s.instanceBase[lf];
                first = pr.firstTriangle;
                box = objectBoxRay(s, lf, r, best.t);
                cur = pr.root;
            }
        }
    }
    if (found) {
        best.u = bestV / bestDet;
        best.v = bestW / bestDet;
    }
    if (kCount && ctr) {
        atomicAdd(&ctr->nodes, nodes);
        atomicAdd(&ctr->triangles, tris);
    }
    return found;
}

// Any hit in [tmin, tmax] over the instance tree.
template <bool kCount>
__device__ bool traceAny2L(const DeviceScene& s, const RayPre& r, float tmin, float tmax, TraversalStack& st,
                           TraversalCounters* ctr)
{
    if (s.nodeCount == 0) return false;
    st.sp = 0;
    std::uint32_t cur = s.tlasRoot;
    bool inBlas = false, hitAny = false;
    int inst = -1;
    BoxRay box = worldBoxRay(r);
    unsigned long long nodes = 0, tris = 0;
    for (;;) {
        if (cur == kInstanceExit) {
            inBlas = false;
            inst = -1;
            box = worldBoxRay(r);
        } else if (cur & kLeafFlag) {
            const std::uint32_t lf = (cur >> 3) & 0x0fffffffU;
            if (!inBlas) {
                st.push(kInstanceExit);
                inBlas = true;
                if (lf >= s.instanceCount) {
                    inst = -1;
                    cur = s.drawnRoot;
                } else {
                    inst = static_cast<int>(lf);
                    box = objectBoxRay(s, lf, r, tmax);
                    cur = s.prototypes[s.instances[lf].prototype].root;
                }
                continue;
            }
            const std::uint32_t count = (cur & 7U) + 1U;
            for (std::uint32_t j = 0; j < count; ++j) {
                float4 a, b, c;
                slotWorld(s, lf + j, inst, a, b, c);
                if (kCount) ++tris;
                float t, vn, wn, det;
                if (intersectTriangle(r, a, b, c, tmin, tmax, t, vn, wn, det)) {
                    hitAny = true;
                    break;
                }
            }
            if (hitAny) break;
        } else {
            if (kCount) ++nodes;
            float tn[4];
            std::uint32_t ref[4];
            testNodeBox(s.nodes, cur, box, tmin, tmax, tn, ref);
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

#undef HDCX_SWAP_IF2

// One- or two-level dispatch, fixed at compile time per kernel instance.
template <bool kCount, bool kTwoLevel>
__device__ __forceinline__ bool traceClosestS(const DeviceScene& s, const RayPre& r, float tmin, float tmax, float loT,
                                              std::uint32_t loPrim, HitRec& best, TraversalStack& st,
                                              TraversalCounters* ctr)
{
    if constexpr (kTwoLevel) return traceClosest2L<kCount>(s, r, tmin, tmax, loT, loPrim, best, st, ctr);
    else return traceClosest<kCount>(s, r, tmin, tmax, loT, loPrim, best, st, ctr);
}
template <bool kCount, bool kTwoLevel>
__device__ __forceinline__ bool traceAnyS(const DeviceScene& s, const RayPre& r, float tmin, float tmax,
                                          TraversalStack& st, TraversalCounters* ctr)
{
    if constexpr (kTwoLevel) return traceAny2L<kCount>(s, r, tmin, tmax, st, ctr);
    else return traceAny<kCount>(s, r, tmin, tmax, st, ctr);
}

// Flattened primitive id -> the scene triangle it reads and its instance
// (-1 for a drawn triangle).
struct TriRef {
    std::uint32_t tri;
    int inst;
};

template <bool kTwoLevel>
__device__ __forceinline__ TriRef resolvePrim(const DeviceScene& s, std::uint32_t prim)
{
    if (!kTwoLevel || prim < s.drawnTriangles) return TriRef{prim, -1};
    std::uint32_t lo = 0, hi = s.instanceCount;  // last instance with base <= prim
    while (hi - lo > 1U) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (s.instanceBase[mid] <= prim) lo = mid;
        else hi = mid;
    }
    const PrototypeRecord pr = s.prototypes[s.instances[lo].prototype];
    return TriRef{pr.firstTriangle + (prim - s.instanceBase[lo]), static_cast<int>(lo)};
}

__device__ __forceinline__ void triangleWorld(const DeviceScene& s, TriRef t, V3& p0, V3& p1, V3& p2)
{
    const std::uint32_t i0 = s.indices[t.tri * 3], i1 = s.indices[t.tri * 3 + 1], i2 = s.indices[t.tri * 3 + 2];
    const float4 a = s.positions[i0], b = s.positions[i1], c = s.positions[i2];
    if (t.inst < 0) {
        p0 = v3(a);
        p1 = v3(b);
        p2 = v3(c);
        return;
    }
    const float* m = s.instances[t.inst].toWorld;
    float w[3];
    TransformPoint(m, a.x, a.y, a.z, w);
    p0 = v3(w[0], w[1], w[2]);
    TransformPoint(m, b.x, b.y, b.z, w);
    p1 = v3(w[0], w[1], w[2]);
    TransformPoint(m, c.x, c.y, c.z, w);
    p2 = v3(w[0], w[1], w[2]);
}

__device__ __forceinline__ float4 cornerNormal(const DeviceScene& s, TriRef t, int corner)
{
    const float4 n = s.normals[t.tri * 3 + corner];
    if (t.inst < 0) return n;
    float w[3];
    TransformNormal(s.instances[t.inst].normal, n.x, n.y, n.z, w);
    return make_float4(w[0], w[1], w[2], 0.0f);
}

// What shading reads of a triangle, from its shading record (one 128-byte
// line, DeviceScene::shading): world-space corners and corner normals
// (through the instance transform, as triangleWorld and cornerNormal), UVs
// and material (an instance's material overrides the triangle's). The same
// values as the scene arrays, read from one line instead of five.
struct TriShading {
    V3 p0, p1, p2;
    V3 n0, n1, n2;
    float2 uv0, uv1, uv2;
    std::uint32_t material;
};
__device__ __forceinline__ TriShading loadTriShading(const DeviceScene& s, TriRef t)
{
    const float4* rec = s.shading + static_cast<std::size_t>(t.tri) * 8U;
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3], r4 = rec[4], r5 = rec[5], r6 = rec[6];
    TriShading o;
    o.p0 = v3(r0.x, r0.y, r0.z);
    o.p1 = v3(r0.w, r1.x, r1.y);
    o.p2 = v3(r1.z, r1.w, r2.x);
    o.n0 = v3(r2.y, r2.z, r2.w);
    o.n1 = v3(r3.x, r3.y, r3.z);
    o.n2 = v3(r3.w, r4.x, r4.y);
    o.uv0 = make_float2(r4.z, r4.w);
    o.uv1 = make_float2(r5.x, r5.y);
    o.uv2 = make_float2(r5.z, r5.w);
    o.material = __float_as_uint(r6.x);
    if (t.inst >= 0) {
        const InstanceRecord& in = s.instances[t.inst];
        float w[3];
        TransformPoint(in.toWorld, o.p0.x, o.p0.y, o.p0.z, w);
        o.p0 = v3(w[0], w[1], w[2]);
        TransformPoint(in.toWorld, o.p1.x, o.p1.y, o.p1.z, w);
        o.p1 = v3(w[0], w[1], w[2]);
        TransformPoint(in.toWorld, o.p2.x, o.p2.y, o.p2.z, w);
        o.p2 = v3(w[0], w[1], w[2]);
        TransformNormal(in.normal, o.n0.x, o.n0.y, o.n0.z, w);
        o.n0 = v3(w[0], w[1], w[2]);
        TransformNormal(in.normal, o.n1.x, o.n1.y, o.n1.z, w);
        o.n1 = v3(w[0], w[1], w[2]);
        TransformNormal(in.normal, o.n2.x, o.n2.y, o.n2.z, w);
        o.n2 = v3(w[0], w[1], w[2]);
        if (in.material != 0xffffffffU) o.material = in.material;
    }
    return o;
}

__device__ __forceinline__ std::uint32_t triangleMaterial(const DeviceScene& s, TriRef t)
{
    if (t.inst >= 0) {
        const std::uint32_t m = s.instances[t.inst].material;
        if (m != 0xffffffffU) return m;
    }
    return s.triMaterial[t.tri];
}

} // namespace hdcx
