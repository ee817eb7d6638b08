// Shading stages of the wavefront path tracer: miss (environment at rays
// that left the scene, hit queue for shade), shade (one surface vertex),
// walk (subsurface random walks deferred by shade).
//
// This file is built with FP32 division and square root at the hardware
// rate (v_rcp_f32-based, within 2.5 ulp, as GLSL allows) instead of the
// correctly rounded expansions: the shading arithmetic is a Monte Carlo
// estimate checked against the FP64 oracle, and the expansions were a
// fifth of the shade kernel's instructions. The stages whose results are
// compared bit for bit with the oracle (ray/triangle hits in extend,
// shadow and the ray queries) live in kernels.hip with IEEE division.
//
// FP contraction is off, as in kernels.hip.
#pragma clang fp contract(off)
#include "kernel_api.h"
#include "dev_wavefront.h"
#include "rcache.h"

#include <type_traits>

namespace hdcx {

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
// miss: the paths whose ray of this bounce left the scene add the
// environment, MIS-weighted against the light pdf of their previous vertex,
// and end (the miss branch of the path loop, vulkan_path_tracer.cpp:
// 1244-1254); the others go to the hit queue (extendQueue[in], whose
// leaders extend has consumed) in queue order, so that shade runs full
// waves of surface hits.
// ---------------------------------------------------------------------------
// One atomic per block of kPairBlock threads (a queue append per wave
// would serialise on the counter: this kernel does little else).
template <class Scenes>
__global__ __launch_bounds__(kPairBlock) void missKernel(Scenes scenes, PathBuffers pb, std::uint32_t in)
{
    __shared__ std::uint32_t waveCount[kPairBlock / 64U], waveBase[kPairBlock / 64U];
    const std::uint32_t count = pb.counters[in * kCounterStride];
    const std::uint32_t* queue = pb.queue[in];
    // Every thread of the block must reach the barriers: iterate over the padded range.
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (count + stride - 1U) / stride * stride;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < padded; i += stride) {
        bool hit = false;
        std::uint32_t pid = 0;
        if (i < count) {
            pid = queue[i];
            hit = __float_as_uint(pb.hit[pid].y) != 0xffffffffU;
            if (!hit) {
                const std::uint32_t flags = pb.flags[pid];
                if (flags & kEnvOnMiss) {
                    const DeviceScene& s = scenes.Get(pid);
                    const float2 mis = pb.mis[pid];
                    const float w = (flags & kPrevDelta) ? 1.0f : powerHeuristic(mis.x, mis.y);
                    pb.radiance[pid] += pb.throughput[pid] * w *
                                        illuminantSpectrum(environmentRgb(s, v3(pb.rayD[pid])), pb.wavelength[pid]);
                }
            }
        }
        pairAppend(hit, pid, pb.extendQueue[in], pb.counters + kHitCounter * kCounterStride, waveCount, waveBase);
    }
}

// First-bounce miss: queue[0] is every lane path of the launch in order,
// three per pixel sample, and the three lanes of a sample hold the same
// camera ray and the same hit (their ray and random state do not depend on
// the lane). One thread handles a sample: the environment's RGB (the dome
// lookup) once, its spectrum at each lane's wavelength, the same sums as
// missKernel per lane; the hit lanes go to the hit queue in path order.
template <class Scenes>
__global__ __launch_bounds__(kPairBlock) void missCameraKernel(Scenes scenes, PathBuffers pb)
{
    __shared__ std::uint32_t waveCount[kPairBlock / 64U], waveBase[kPairBlock / 64U];
    const std::uint32_t samples = pb.counters[0] / 3U;
    const std::uint32_t stride = gridDim.x * blockDim.x;
    const std::uint32_t padded = (samples + stride - 1U) / stride * stride;
    for (std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < padded; j += stride) {
        bool hit = false;
        const std::uint32_t pid0 = 3U * j;
        if (j < samples) {
            hit = __float_as_uint(pb.hit[pid0].y) != 0xffffffffU;
            if (!hit) {
                const DeviceScene& s = scenes.Get(pid0);
                bool have = false;
                V3 rgb = v3(0.0f);
                for (std::uint32_t lane = 0; lane < 3U; ++lane) {
                    const std::uint32_t pid = pid0 + lane;
                    const std::uint32_t flags = pb.flags[pid];
                    if (!(flags & kEnvOnMiss)) continue;
                    if (!have) {
                        rgb = environmentRgb(s, v3(pb.rayD[pid0]));
                        have = true;
                    }
                    float w = 1.0f;
                    if (!(flags & kPrevDelta)) {
                        const float2 mis = pb.mis[pid];
                        w = powerHeuristic(mis.x, mis.y);
                    }
                    pb.radiance[pid] += pb.throughput[pid] * w * illuminantSpectrum(rgb, pb.wavelength[pid]);
                }
            }
        }
        pairAppend3(hit, pid0, pb.extendQueue[0], pb.counters + kHitCounter * kCounterStride, waveCount, waveBase);
    }
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
//
// kShadeSimple is the variant for scenes whose materials have no coat,
// sheen, transmission, subsurface or translucency (DeviceScene::
// simpleMaterials, checked on every material at upload): those lobes are
// compiled out. What it skips is exactly what evaluates to a no-op for such
// materials (zero weights, zero selection probabilities, factors of 1), and
// it draws the same random numbers (the transmission pick is drawn and
// ignored), so its images equal the full kernel's.
// ---------------------------------------------------------------------------
// Occupancy: the simple variant and the one-scene full variants are held to
// 3 waves per SIMD (168 VGPRs, a few spilled; at the 188-252 registers they
// need otherwise the kernel runs 2 waves): C1 shade 16.5 -> 15.2 ms per
// frame, the 10,000-instance frame's 560 -> 520 ms; 4 waves were slower
// (16.4 ms). The full variants over a scene table (the batch renderer) keep
// their registers (3 waves measured 0.5% slower there).
template <class Scenes, bool k2L, int kWalk>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    (kWalk == kShadeSimple || std::is_same<Scenes, SingleScene>::value) ? 3 : 1))) void shadeKernel(Scenes scenes, PathBuffers pb, std::uint32_t in, int bounce,
                                                     int lastBounce, CacheParams cp, int share)
{
    __shared__ std::uint32_t stackMem[kLdsStack * kBlock];
    HDCX_TRAVERSAL_STACK(st, stackMem);
    // Paths whose ray hit a triangle (missKernel queued them) or whose
    // deferred walk is to be resumed.
    const std::uint32_t count =
        kWalk == kShadeWalkResume ? pb.counters[3 * kCounterStride] : pb.counters[kHitCounter * kCounterStride];
    const std::uint32_t* queue = kWalk == kShadeWalkResume ? pb.walkQueue : pb.extendQueue[in];
    constexpr bool kFull = kWalk != kShadeSimple;
    if (kWalk != kShadeWalkResume && blockIdx.x == 0 && threadIdx.x == 0) pb.totals[4] += count;
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
        // The continuing ray (and, with cut-outs, random state) compared by
        // the lane pairing below.
        std::uint32_t pairWords[7] = {0U, 0U, 0U, 0U, 0U, 0U, 0U};
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
            // The medium matters only under kInMedium, which only the full
            // variant's transmission lobe sets (with the medium).
            float4 med = kFull ? pb.medium[pid] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float2 spreadOut = make_float2(0.0f, 0.0f);
            bool spreadSet = false;

            {
                const float distance = hit4.x, bu = hit4.z, bv = hit4.w;
                bool mediumScattered = false;
                if (kFull && (flags & kInMedium)) {
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
                    const TriShading ts = loadTriShading(s, tr);
                    const V3 p0 = ts.p0, p1 = ts.p1, p2 = ts.p2;
                    const V3 gn = normalize(cross(p1 - p0, p2 - p0));
                    const bool front = dot(gn, rd) < 0.0f;
                    const V3 ogn = front ? gn : -gn;
                    const float w0 = 1.0f - bu - bv;
                    V3 normal;
                    {
                        V3 an = ts.n0 * w0 + ts.n1 * bu + ts.n2 * bv;
                        if (dot(an, an) < 1e-12f) an = gn;
                        else {
                            an = normalize(an);
                            if (dot(an, gn) < 0.0f) an = -an;
                        }
                        normal = front ? an : -an;
                    }
                    const float2 t0 = ts.uv0, t1 = ts.uv1, t2 = ts.uv2;
                    const float su = t0.x * w0 + t1.x * bu + t2.x * bv, sv = t0.y * w0 + t1.y * bu + t2.y * bv;
                    const V3 hp = ro + rd * distance;
                    // The record by value: the fields this vertex reads are
                    // loaded together here, instead of one dependent load
                    // after every texture-slot test below.
                    const GpuMaterialRecord m = s.materials[ts.material];

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

                    c.transmission = kFull ? clampf(m.transmissionOpacityIor.x, 0.0f, 1.0f) : 0.0f;
                    V3 transRgb = v3(m.transmissionColorThinWalled);
                    if (c.transmission > 0.0f) transRgb = matRgb(s, m.textureIndices1.z, su, sv, transRgb);
                    const float transmissionColor = reflectanceSpectrum(transRgb, wl);
                    c.subsurface = 0.0f;
                    c.diffuseColor = baseColor;
                    V3 sssColorRgb = v3(m.subsurfaceColor);
                    V3 sssRadiusRgb = vmax(v3(m.subsurfaceRadius), 0.001f);
                    if (kFull && (m.subsurfaceWeightScale.x > 0.0f || m.textureIndices2.x != kMissingTexture)) {
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
                    if (kFull && (m.coatWeightRoughnessIor.x > 0.0f || m.textureIndices3.z != kMissingTexture)) {
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
                    if (kFull) {
                        c.sheen = clampf(matScalar(s, m.textureIndices1.w, su, sv, m.coatColor.w), 0.0f, 1.0f);
                        c.sheenColor = reflectanceSpectrum(matRgb(s, m.textureIndices2.w, su, sv, v3(m.sheenColorMode)), wl);
                        c.sheenRoughness = clampf(matScalar(s, m.textureIndices4.y, su, sv, m.transmissionScatter.w), 0.01f, 1.0f);
                        c.sheenZeltner = m.sheenColorMode.w > 0.5f;
                        c.translucentWeight = clampf(matScalar(s, m.textureIndices5.x, su, sv, m.translucentColorWeight.w), 0.0f, 1.0f);
                        c.translucentColor = reflectanceSpectrum(matRgb(s, m.textureIndices5.y, su, sv, v3(m.translucentColorWeight)), wl);
                    } else {
                        c.sheen = 0.0f;
                        c.sheenColor = 0.0f;
                        c.sheenRoughness = 0.01f;
                        c.sheenZeltner = false;
                        c.translucentWeight = 0.0f;
                        c.translucentColor = 0.0f;
                    }

                    // Lobe probabilities (:1435-1464).
                    c.view = normalize(-rd);
                    const float nv = fmaxf(dot(normal, c.view), 0.0f);
                    c.sheenAlbedo = !kFull ? 0.0f
                                    : c.sheenZeltner ? sheenAlbedoZeltner(nv, c.sheenRoughness) : sheenAlbedoCK(nv, c.sheenRoughness);
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
                            const float direct = evalDirect<kFull>(c, ld, L.controls.x, L.controls.y);
                            if (direct > 0.0f) {
                                const float w = L.textureInfo.z == 0U
                                    ? powerHeuristic(environmentLightPdf(s, hp, normal, ld, c.diffE, dSpecE, c.translucentE, lightTotal),
                                                     surfacePdf<kFull>(c, ld))
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
                        const float direct = evalDirect<kFull>(c, sun, 1.0f, 1.0f);
                        if (direct > 0.0f) {
                            const float contribution = throughput * direct * illuminantSpectrum(v3(3.45575f, 3.14159f, 2.82743f), wl);
                            queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, sun), sun, 10000.0f, contribution, 0U);
                        }
                        const V3 sky = cosineHemisphere(normal, rng);
                        const float skyDirect = evalDirect<kFull>(c, sky, 1.0f, 0.0f);
                        if (skyDirect > 0.0f) {
                            const float contribution = throughput * skyDirect * illuminantSpectrum(environmentRgb(s, sky), wl) *
                                                       (kPi / fmaxf(dot(normal, sky), 1e-4f));
                            queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, sky), sky, 10000.0f, contribution, 0U);
                        }
                        if (c.translucentWeight > 0.0f) {
                            const V3 back = cosineHemisphere(-normal, rng);
                            const float backDirect = evalDirect<kFull>(c, back, 1.0f, 0.0f);
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
                                const float direct = evalDirect<kFull>(c, ld, 1.0f, 1.0f);
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
                                    const float w = powerHeuristic(lightPdf, surfacePdf<kFull>(c, ld));
                                    const float contribution = throughput * direct * illuminantSpectrum(le, wl) * (w / lightPdf);
                                    queueShadow(pb, pid, slot, offsetOrigin(hp, ogn, ld), ld, fmaxf(dist - 0.003f, 0.001f),
                                                contribution, kShadowMeshLight);
                                }
                            }
                        }
                    }

                    if (!lastBounce && !cacheStop) {
                        bool walked = false;
                        const float transmissionPick = randomFloat(rng);
                        if (kFull && transmissionPick < c.transmission) {
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
                            if (kFull && c.pSheen > 0.0f && randomFloat(rng) < c.pSheen) {
                                rd = cosineHemisphere(normal, rng);
                                const float sp = fmaxf(dot(normal, rd), 0.0f) / kPi;
                                throughput *= c.sheen * c.sheenColor * sheenResponse(normal, c.view, rd, c.sheenRoughness, c.sheenZeltner) /
                                              fmaxf(c.pSheen * sp, 1e-6f);
                                flags |= kEnvOnMiss;
                            } else {
                                if (kFull) throughput *= (1.0f - c.sheen * c.sheenAlbedo) / fmaxf(1.0f - c.pSheen, 1e-4f);
                                if (kFull && c.pCoat > 0.0f && randomFloat(rng) < c.pCoat) {
                                    const V3 inc = rd;
                                    rd = sampleGgx(rd, normal, c.tangent, c.crx, c.cry, rng);
                                    throughput *= c.coat * ggxSampleWeight(inc, rd, normal, c.tangent, c.crx, c.cry, c.coatF0) / c.pCoat;
                                    if (fmaxf(c.crx, c.cry) < 0.2f) flags |= kEnvOnMiss;
                                } else {
                                    if (kFull) throughput *= (1.0f - c.coatContribution) / fmaxf(1.0f - c.pCoat, 1e-4f);
                                    if (randomFloat(rng) < c.pSpec) {
                                        const V3 inc = rd;
                                        rd = sampleGgx(rd, normal, c.tangent, c.rx, c.ry, rng);
                                        throughput *= ggxSampleWeight(inc, rd, normal, c.tangent, c.rx, c.ry, c.f0) / c.pSpec;
                                        if (fmaxf(c.rx, c.ry) < 0.2f) flags |= kEnvOnMiss;
                                    } else if (kFull && c.pTranslucent > 0.0f && randomFloat(rng) < c.pTranslucent) {
                                        throughput *= c.translucentContribution / fmaxf((1.0f - c.pSpec) * c.pTranslucent, 1e-4f);
                                        rd = cosineHemisphere(-normal, rng);
                                        ro = offsetOrigin(hp, ogn, rd);
                                    } else {
                                        throughput *= c.diffuseContribution / fmaxf((1.0f - c.pSpec) * (1.0f - c.pTranslucent), 1e-4f);
                                        if (kFull && c.subsurface > 0.0f && randomFloat(rng) < c.subsurface) {
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
                                    mis.x = surfacePdf<kFull>(c, rd);
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
                // The shadow slots end at the first light of 0xffffffff: a
                // path that queued one or two ends its list here (a path
                // with none is not in the shadow queue, and no pairing or
                // shadow pass reads its slots).
                if (slot > 0 && slot < kShadowSlots) pb.shadowLight[static_cast<std::size_t>(slot) * pb.capacity + pid] = 0xffffffffU;
                if (!deferred) {
                    if (keep) {
                        pb.rayO[pid] = make_float4(ro.x, ro.y, ro.z, 0.0f);
                        pb.rayD[pid] = make_float4(rd.x, rd.y, rd.z, 0.0f);
                        pb.flags[pid] = flags;
                        pb.mis[pid] = mis;
                        if (kFull) pb.medium[pid] = med;
                        pairWords[0] = __float_as_uint(ro.x);
                        pairWords[1] = __float_as_uint(ro.y);
                        pairWords[2] = __float_as_uint(ro.z);
                        pairWords[3] = __float_as_uint(rd.x);
                        pairWords[4] = __float_as_uint(rd.y);
                        pairWords[5] = __float_as_uint(rd.z);
                        if (!s.opaque) pairWords[6] = rng;
                    }
                    pb.rng[pid] = rng;
                    pb.throughput[pid] = throughput;
                    pb.radiance[pid] = radiance;
                    if (spreadSet) pb.spread[pid] = spreadOut;
                }
            }
        }
        waveAppend(keep, pid, outQueue, outCounter);
        // Lane sharing (share != 0): paths whose next ray equals, bit for
        // bit, that of an earlier lane of their pixel sample held by one of
        // the two preceding lanes of this wave do not trace it; the tracing
        // paths go to the next extend's queue with the mask of their
        // followers (pairLanes). The lanes of a sample have consecutive path
        // ids and sit next to each other in the hit queue whenever they hit.
        if (share && !lastBounce) {
            std::uint32_t followers;
            const bool traces = pairLanes<7>(keep, pid, pairWords, followers);
            if (traces) pb.follow[pid] = followers;
            waveAppend(traces, pid, pb.extendQueue[in ^ 1U], pb.counters + (4U + (in ^ 1U)) * kCounterStride);
        }
        if constexpr (kWalk != kShadeWalkResume) {
            // A resumed path queued its shadow rays when it was deferred.
            waveAppend(slot > 0, pid, pb.shadowQueue, pb.counters + 2 * kCounterStride);
            if (share) {
                // The same grouping of the shadow rays just queued: every
                // slot's light, origin, extent and direction must be equal.
                // The shadow kernel traces each group once and applies the
                // visibility to every lane of it with the lane's own
                // wavelength and contribution.
                std::uint32_t w[8 * kShadowSlots];
                for (int j = 0; j < 8 * kShadowSlots; ++j) w[j] = 0U;
                if (slot > 0) {
                    for (int k = 0; k < kShadowSlots; ++k) {
                        const std::size_t idx = static_cast<std::size_t>(k) * pb.capacity + pid;
                        const std::uint32_t light = k < slot ? pb.shadowLight[idx] : 0xffffffffU;
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
                const bool traces = pairLanes<8 * kShadowSlots>(slot > 0, pid, w, followers);
                if (traces) pb.shadowFollow[pid] = followers;
                waveAppend(traces, pid, pb.shadowLeaders, pb.counters + 6U * kCounterStride);
            }
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
    HDCX_TRAVERSAL_STACK(st, stackMem);
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
// Host launchers.
// ---------------------------------------------------------------------------
template <class Scenes>
hipError_t LaunchMissT(const Scenes& s, const PathBuffers& pb, std::uint32_t in, std::uint32_t maxWork, unsigned maxBlocks,
                       hipStream_t st)
{
    const unsigned blocks = StageGrid(missKernel<Scenes>, maxWork, maxBlocks, kPairBlock);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(missKernel<Scenes>), dim3(blocks), dim3(kPairBlock), 0, st, s, pb, in);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchShadeT(const Scenes& s, const PathBuffers& pb, std::uint32_t in, int bounce, int last,
                        const CacheParams& cp, int walk, int share, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    if (walk == kShadeSimple)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeSimple>),
                           dim3(StageGrid(shadeKernel<Scenes, k2L, kShadeSimple>, maxWork, maxBlocks)), dim3(kBlock), 0, st, s,
                           pb, in, bounce, last, cp, share);
    else if (walk == kShadeWalkDefer)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkDefer>),
                           dim3(StageGrid(shadeKernel<Scenes, k2L, kShadeWalkDefer>, maxWork, maxBlocks)), dim3(kBlock), 0, st,
                           s, pb, in, bounce, last, cp, share);
    else if (walk == kShadeWalkResume)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkResume>),
                           dim3(StageGrid(shadeKernel<Scenes, k2L, kShadeWalkResume>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, in, bounce, last, cp, share);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(shadeKernel<Scenes, k2L, kShadeWalkInline>),
                           dim3(StageGrid(shadeKernel<Scenes, k2L, kShadeWalkInline>, maxWork, maxBlocks)), dim3(kBlock), 0,
                           st, s, pb, in, bounce, last, cp, share);
    return hipGetLastError();
}
template <class Scenes, bool k2L>
hipError_t LaunchWalkT(const Scenes& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    hipLaunchKernelGGL(HIP_KERNEL_NAME(walkKernel<Scenes, k2L>), dim3(StageGrid(walkKernel<Scenes, k2L>, maxWork, maxBlocks)),
                       dim3(kBlock), 0, st, s,
                       pb);
    return hipGetLastError();
}
hipError_t LaunchMissCamera(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st)
{
    if (maxWork % 3U != 0U) return hipErrorInvalidValue;  // not three lanes per pixel sample
    const unsigned blocks = StageGrid(missCameraKernel<SingleScene>, maxWork / 3U, maxBlocks, kPairBlock);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(missCameraKernel<SingleScene>), dim3(blocks), dim3(kPairBlock), 0, st,
                       SingleScene{s}, pb);
    return hipGetLastError();
}
hipError_t LaunchMiss(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, std::uint32_t maxWork,
                      unsigned maxBlocks, hipStream_t st)
{
    return LaunchMissT<SingleScene>(SingleScene{s}, pb, in, maxWork, maxBlocks, st);
}
hipError_t LaunchShade(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int bounce, int last,
                       const CacheParams& cp, int walk, int share, std::uint32_t maxWork, unsigned maxBlocks, hipStream_t st)
{
    return s.instanceCount
        ? LaunchShadeT<SingleScene, true>(SingleScene{s}, pb, in, bounce, last, cp, walk, share, maxWork, maxBlocks, st)
        : LaunchShadeT<SingleScene, false>(SingleScene{s}, pb, in, bounce, last, cp, walk, share, maxWork, maxBlocks, st);
}
hipError_t LaunchWalk(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                      hipStream_t st)
{
    return s.instanceCount ? LaunchWalkT<SingleScene, true>(SingleScene{s}, pb, maxWork, maxBlocks, st)
                           : LaunchWalkT<SingleScene, false>(SingleScene{s}, pb, maxWork, maxBlocks, st);
}
hipError_t LaunchMissCameraBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                                 hipStream_t st)
{
    if (maxWork % 3U != 0U) return hipErrorInvalidValue;  // not three lanes per pixel sample
    const unsigned blocks = StageGrid(missCameraKernel<SceneTable>, maxWork / 3U, maxBlocks, kPairBlock);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(missCameraKernel<SceneTable>), dim3(blocks), dim3(kPairBlock), 0, st,
                       SceneTable{table, pb.pathScene}, pb);
    return hipGetLastError();
}
hipError_t LaunchMissBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t in, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st)
{
    return LaunchMissT<SceneTable>(SceneTable{table, pb.pathScene}, pb, in, maxWork, maxBlocks, st);
}
hipError_t LaunchShadeBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t in, int bounce,
                            int last, const CacheParams& cp, int walk, int share, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchShadeT<SceneTable, true>(t, pb, in, bounce, last, cp, walk, share, maxWork, maxBlocks, st)
                    : LaunchShadeT<SceneTable, false>(t, pb, in, bounce, last, cp, walk, share, maxWork, maxBlocks, st);
}
hipError_t LaunchWalkBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st)
{
    const SceneTable t{table, pb.pathScene};
    return twoLevel ? LaunchWalkT<SceneTable, true>(t, pb, maxWork, maxBlocks, st)
                    : LaunchWalkT<SceneTable, false>(t, pb, maxWork, maxBlocks, st);
}

} // namespace hdcx
