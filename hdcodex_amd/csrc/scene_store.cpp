// Scene upload (validation, packed arrays, BVH build or refit, dome
// importance tables) and wavefront path storage; see scene_store.h.
#include "scene_store.h"
#include "dev_traverse.h"
#include "lbvh_host.hpp"
#include "records.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace hdcx {
namespace {

void ValidateTextureRef(std::uint32_t ti, std::uint32_t textureCount, const char* what)
{
    if (ti == kMissingTexture) return;
    if (ti & kUdimTextureFlag) {
        const std::uint32_t base = ti & 0xffU;
        const std::uint32_t tiles = static_cast<std::uint32_t>(__builtin_popcount((ti >> 8U) & kUdimTileMask));
        if (base + tiles > textureCount) throw Failure(HDCX_ERR_INVALID_SCENE, std::string(what) + ": UDIM set out of range");
        return;
    }
    if (ti >= textureCount) throw Failure(HDCX_ERR_INVALID_SCENE, std::string(what) + ": texture index out of range");
}

} // namespace

void ValidateScene(const hdcx_scene_view& v)
{
    if (v.triangle_count && (!v.indices || !v.triangle_materials || !v.texcoords || !v.normals || !v.emissive))
        throw Failure(HDCX_ERR_INVALID_SCENE, "triangle arrays missing");
    if (v.vertex_count && !v.positions) throw Failure(HDCX_ERR_INVALID_SCENE, "positions missing");
    if (v.material_count == 0 || !v.materials) throw Failure(HDCX_ERR_INVALID_SCENE, "no materials");
    if (v.light_count > HDCX_MAX_LIGHTS) throw Failure(HDCX_ERR_INVALID_SCENE, "more than 64 lights");
    if (!v.lights) throw Failure(HDCX_ERR_INVALID_SCENE, "light table missing");
    if (v.texture_count > HDCX_MAX_TEXTURES) throw Failure(HDCX_ERR_INVALID_SCENE, "more than 256 textures");
    if (v.texture_count && !v.textures) throw Failure(HDCX_ERR_INVALID_SCENE, "texture views missing");
    for (std::uint64_t i = 0; i < static_cast<std::uint64_t>(v.vertex_count) * 4; ++i)
        if (!std::isfinite(v.positions[i])) throw Failure(HDCX_ERR_INVALID_SCENE, "non-finite vertex position");
    for (std::uint64_t i = 0; i < static_cast<std::uint64_t>(v.triangle_count) * 3; ++i)
        if (v.indices[i] >= v.vertex_count) throw Failure(HDCX_ERR_INVALID_SCENE, "vertex index out of range");
    for (std::uint32_t t = 0; t < v.triangle_count; ++t)
        if (v.triangle_materials[t] >= v.material_count) throw Failure(HDCX_ERR_INVALID_SCENE, "material index out of range");
    for (std::uint32_t i = 0; i < v.texture_count; ++i) {
        const hdcx_texture_view& t = v.textures[i];
        if (t.width == 0 || t.height == 0 || !t.texels || t.format > HDCX_TEXTURE_RGBA32F)
            throw Failure(HDCX_ERR_INVALID_SCENE, "invalid texture");
    }
    const auto* mats = static_cast<const GpuMaterialRecord*>(v.materials);
    for (std::uint32_t m = 0; m < v.material_count; ++m) {
        const U4* slots[6] = {&mats[m].textureIndices0, &mats[m].textureIndices1, &mats[m].textureIndices2,
                              &mats[m].textureIndices3, &mats[m].textureIndices4, &mats[m].textureIndices5};
        for (const U4* u : slots) {
            ValidateTextureRef(u->x, v.texture_count, "material");
            ValidateTextureRef(u->y, v.texture_count, "material");
            ValidateTextureRef(u->z, v.texture_count, "material");
            ValidateTextureRef(u->w, v.texture_count, "material");
        }
    }
    const auto* ls = static_cast<const GpuLightRecord*>(v.lights);
    for (std::uint32_t l = 0; l < v.light_count; ++l) {
        ValidateTextureRef(ls[l].textureInfo.x, v.texture_count, "light");
        if (ls[l].textureInfo.w != kMissingTexture) {
            if (ls[l].textureInfo.w >= v.texture_count ||
                v.textures[ls[l].textureInfo.w].format != HDCX_TEXTURE_RGBA32F)
                throw Failure(HDCX_ERR_INVALID_SCENE, "light importance texture invalid");
        }
    }
}

namespace {

void UploadHostBvh(SceneStore* t, const Bvh4& bvh, int builder, hipStream_t st)
{
    if (bvh.maxStack > static_cast<std::uint32_t>(kMaxStack))
        throw Failure(HDCX_ERR_INVALID_SCENE, "BVH deeper than the traversal stack");
    t->nodes.Upload(bvh.nodes.data(), bvh.nodes.size() * sizeof(Bvh4Node), st);
    t->tris.Upload(bvh.triangles.data(), bvh.triangles.size() * sizeof(float), st);
    MakeRefitPlan(bvh.nodes, st, t->refitPlan);
    t->bvhInfo = hdcx_bvh_info{};
    t->bvhInfo.node_count = static_cast<std::uint32_t>(bvh.nodes.size());
    t->bvhInfo.slot_count = static_cast<std::uint32_t>(bvh.triangles.size() / 12);
    t->bvhInfo.max_stack = bvh.maxStack;
    t->bvhInfo.depth = bvh.depth;
    t->bvhInfo.builder = builder;
}

// Builds the BVH of the uploaded positions / indices with the configured
// builder. The device builder falls back to the host one for a single
// triangle or a tree deeper than the traversal stack.
void BuildSceneBvh(SceneStore* t, const hdcx_scene_view& v, const float lo[3], const float hi[3], hipStream_t st)
{
    const std::uint32_t n = v.triangle_count;
    int builder = t->bvhOptions.builder;
    if (builder != HDCX_BVH_HOST_SAH && n < 2) builder = HDCX_BVH_HOST_SAH;
    if (builder == HDCX_BVH_DEVICE_LBVH) {
        DeviceBvhStats bs;
        t->lbvh.Build(t->positions.as<float4>(), t->indices.as<std::uint32_t>(), n, lo, hi, t->bvhOptions.top_clusters,
                      st, t->nodes, t->tris, t->refitPlan, bs);
        if (bs.maxStack <= static_cast<std::uint32_t>(kMaxStack)) {
            t->bvhInfo = hdcx_bvh_info{};
            t->bvhInfo.node_count = bs.nodeCount;
            t->bvhInfo.slot_count = n;
            t->bvhInfo.max_stack = bs.maxStack;
            t->bvhInfo.depth = bs.depth;
            t->bvhInfo.builder = HDCX_BVH_DEVICE_LBVH;
            t->bvhInfo.top_clusters = bs.topClusters;
            return;
        }
        builder = HDCX_BVH_HOST_SAH;
    }
    if (builder == HDCX_BVH_SERIAL_LBVH) {
        LbvhBuildInfo info;
        Bvh4 bvh = BuildBvh4LbvhSerial(v.positions, v.indices, n, t->bvhOptions.top_clusters, &info);
        ValidateBvh4(bvh, n);
        if (bvh.maxStack <= static_cast<std::uint32_t>(kMaxStack)) {
            UploadHostBvh(t, bvh, HDCX_BVH_SERIAL_LBVH, st);
            t->bvhInfo.top_clusters = info.topClusters;
            return;
        }
    }
    Bvh4 bvh = BuildBvh4Host(v.positions, v.indices, n);
    ValidateBvh4(bvh, n);
    UploadHostBvh(t, bvh, HDCX_BVH_HOST_SAH, st);
}

} // namespace

void UploadScene(SceneStore* t, const hdcx_scene_view* view, hipStream_t st)
{
    const bool hadGeometry = t->geometry;
    t->geometry = false;
    if (!view || view->vertex_count == 0 || view->triangle_count == 0) {
        t->builtIndices.clear();
        t->builtVertexCount = 0;
        return;
    }
    ValidateScene(*view);
    const hdcx_scene_view& v = *view;

    float lo[3], hi[3];
    VertexBounds(v.positions, v.vertex_count, lo, hi);
    for (int k = 0; k < 3; ++k) {
        t->sceneLo[k] = lo[k];
        t->sceneInvExtent[k] = hi[k] > lo[k] ? 1.0F / (hi[k] - lo[k]) : 1.0F;
    }
    const bool refit = t->bvhOptions.refit != 0U && hadGeometry && t->builtVertexCount == v.vertex_count &&
                       t->builtIndices.size() == static_cast<std::size_t>(v.triangle_count) * 3 &&
                       std::memcmp(t->builtIndices.data(), v.indices, t->builtIndices.size() * 4) == 0;
    t->positions.Upload(v.positions, static_cast<std::size_t>(v.vertex_count) * 16, st);
    t->indices.Upload(v.indices, static_cast<std::size_t>(v.triangle_count) * 12, st);
    const auto buildStart = std::chrono::steady_clock::now();
    if (refit) {
        RefitBvh4(t->positions.as<float4>(), t->indices.as<std::uint32_t>(), t->bvhInfo.slot_count,
                  t->nodes.as<Bvh4Node>(), t->tris.as<float4>(), t->refitPlan, st);
        ++t->bvhInfo.refits;
    } else {
        BuildSceneBvh(t, v, lo, hi, st);
        t->builtIndices.assign(v.indices, v.indices + static_cast<std::size_t>(v.triangle_count) * 3);
        t->builtVertexCount = v.vertex_count;
    }
    t->bvhInfo.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - buildStart).count();
    t->triMaterial.Upload(v.triangle_materials, static_cast<std::size_t>(v.triangle_count) * 4, st);
    t->texcoords.Upload(v.texcoords, static_cast<std::size_t>(v.triangle_count) * 24, st);
    t->normals.Upload(v.normals, static_cast<std::size_t>(v.triangle_count) * 48, st);
    t->emissive.Upload(v.emissive, static_cast<std::size_t>(v.triangle_count) * 16, st);
    t->materials.Upload(v.materials, static_cast<std::size_t>(v.material_count) * HDCX_MATERIAL_BYTES, st);
    const std::uint32_t lightRecords = std::max<std::uint32_t>(v.light_count, 1);
    t->lights.Upload(v.lights, static_cast<std::size_t>(lightRecords) * HDCX_LIGHT_BYTES, st);

    // Texel pool (16-byte aligned entries).
    std::vector<TexDesc> descs(std::max<std::uint32_t>(v.texture_count, 1));
    std::size_t poolBytes = 0;
    for (std::uint32_t i = 0; i < v.texture_count; ++i) {
        const hdcx_texture_view& tv = v.textures[i];
        const std::size_t bytes = static_cast<std::size_t>(tv.width) * tv.height * (tv.format == HDCX_TEXTURE_RGBA32F ? 16 : 4);
        descs[i] = TexDesc{poolBytes, tv.width, tv.height, tv.format, 0};
        poolBytes += (bytes + 15) & ~static_cast<std::size_t>(15);
    }
    std::vector<unsigned char> pool(std::max<std::size_t>(poolBytes, 16));
    for (std::uint32_t i = 0; i < v.texture_count; ++i) {
        const hdcx_texture_view& tv = v.textures[i];
        const std::size_t bytes = static_cast<std::size_t>(tv.width) * tv.height * (tv.format == HDCX_TEXTURE_RGBA32F ? 16 : 4);
        std::memcpy(pool.data() + descs[i].offset, tv.texels, bytes);
    }
    t->textures.Upload(descs.data(), descs.size() * sizeof(TexDesc), st);
    t->texels.Upload(pool.data(), pool.size(), st);

    // Dome importance tables re-laid out per light: marginal CDF column,
    // conditional CDF and pdf planes of the packed importance texture.
    const auto* ls = static_cast<const GpuLightRecord*>(v.lights);
    std::vector<ImportanceDesc> imp(lightRecords, ImportanceDesc{0, 0, 0, 0, 0});
    std::vector<float> impPool;
    for (std::uint32_t l = 0; l < v.light_count; ++l) {
        const std::uint32_t ti = ls[l].textureInfo.w;
        if (ls[l].textureInfo.z != 0U || ti == kMissingTexture) continue;
        const hdcx_texture_view& tv = v.textures[ti];
        const float* px = static_cast<const float*>(tv.texels);
        ImportanceDesc d{};
        d.width = tv.width;
        d.height = tv.height;
        d.marginal = impPool.size();
        for (std::uint32_t y = 0; y < tv.height; ++y) impPool.push_back(px[(static_cast<std::size_t>(y) * tv.width) * 4 + 1]);
        d.conditional = impPool.size();
        for (std::size_t p = 0; p < static_cast<std::size_t>(tv.width) * tv.height; ++p) impPool.push_back(px[p * 4 + 0]);
        d.pdf = impPool.size();
        for (std::size_t p = 0; p < static_cast<std::size_t>(tv.width) * tv.height; ++p) impPool.push_back(px[p * 4 + 2]);
        imp[l] = d;
    }
    if (impPool.empty()) impPool.push_back(0.0F);
    t->importance.Upload(imp.data(), imp.size() * sizeof(ImportanceDesc), st);
    t->importancePool.Upload(impPool.data(), impPool.size() * sizeof(float), st);
    Check(hipStreamSynchronize(st), "scene upload");

    DeviceScene& s = t->scene;
    s.nodes = t->nodes.as<Bvh4Node>();
    s.tris = t->tris.as<float4>();
    s.positions = t->positions.as<float4>();
    s.indices = t->indices.as<std::uint32_t>();
    s.triMaterial = t->triMaterial.as<std::uint32_t>();
    s.texcoords = t->texcoords.as<float2>();
    s.normals = t->normals.as<float4>();
    s.emissive = t->emissive.as<float4>();
    s.materials = t->materials.as<GpuMaterialRecord>();
    s.lights = t->lights.as<GpuLightRecord>();
    s.textures = t->textures.as<TexDesc>();
    s.texels = t->texels.as<unsigned char>();
    s.importance = t->importance.as<ImportanceDesc>();
    s.importancePool = t->importancePool.as<float>();
    s.primitiveCount = v.triangle_count;
    s.lightCount = v.light_count;
    s.opaque = v.opaque ? 1U : 0U;
    s.hasEmissive = v.emissive[(static_cast<std::size_t>(v.triangle_count) - 1) * 4] > 0.0F ? 1U : 0U;
    s.textureCount = v.texture_count;
    s.nodeCount = t->bvhInfo.node_count;
    t->geometry = true;
}

void PathPool::EnsurePaths(std::uint32_t paths)
{
    if (!counters.bytes()) {
        counters.Allocate(64);
        totals.Allocate(64);
        Check(hipMemset(totals.as<void>(), 0, 64), "memset totals");
    }
    if (paths <= capacity) return;
    const std::size_t n = paths;
    rayO.Allocate(n * 16);
    rayD.Allocate(n * 16);
    hit.Allocate(n * 16);
    rng.Allocate(n * 4);
    wavelength.Allocate(n * 4);
    throughput.Allocate(n * 4);
    radiance.Allocate(n * 4);
    flags.Allocate(n * 4);
    mis.Allocate(n * 8);
    medium.Allocate(n * 16);
    shadowO.Allocate(n * 16 * kShadowSlots);
    shadowD.Allocate(n * 16 * kShadowSlots);
    shadowLight.Allocate(n * 4 * kShadowSlots);
    queue0.Allocate(n * 4);
    queue1.Allocate(n * 4);
    shadowQueue.Allocate(n * 4);
    lambda0.Allocate(n * 4);
    spread.Allocate(n * 8);
    pathScene.Allocate(n * 4);
    capacity = paths;
}

PathBuffers PathPool::Paths()
{
    PathBuffers pb{};
    pb.rayO = rayO.as<float4>();
    pb.rayD = rayD.as<float4>();
    pb.hit = hit.as<float4>();
    pb.rng = rng.as<std::uint32_t>();
    pb.wavelength = wavelength.as<float>();
    pb.throughput = throughput.as<float>();
    pb.radiance = radiance.as<float>();
    pb.flags = flags.as<std::uint32_t>();
    pb.mis = mis.as<float2>();
    pb.medium = medium.as<float4>();
    pb.shadowO = shadowO.as<float4>();
    pb.shadowD = shadowD.as<float4>();
    pb.shadowLight = shadowLight.as<std::uint32_t>();
    pb.queue[0] = queue0.as<std::uint32_t>();
    pb.queue[1] = queue1.as<std::uint32_t>();
    pb.shadowQueue = shadowQueue.as<std::uint32_t>();
    pb.counters = counters.as<std::uint32_t>();
    pb.totals = totals.as<unsigned long long>();
    pb.lambda0 = lambda0.as<float>();
    pb.spread = spread.as<float2>();
    pb.pathScene = pathScene.as<std::uint32_t>();
    pb.capacity = capacity;
    return pb;
}

} // namespace hdcx
