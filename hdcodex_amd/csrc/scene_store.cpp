// Scene upload (validation, packed arrays, BVH build or refit, dome
// importance tables) and wavefront path storage; see scene_store.h.
#include "scene_store.h"
#include "dev_traverse.h"
#include "instance_math.h"
#include "lbvh_host.hpp"
#include "records.h"
#include "tables.h"

#include <algorithm>
#include <chrono>
#include <cstddef>
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
    if (v.triangle_count && (!v.indices || !v.triangle_materials || !v.texcoords || !v.normals))
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
    if (v.prototype_count) {
        if (!v.prototypes) throw Failure(HDCX_ERR_INVALID_SCENE, "prototype table missing");
        std::uint64_t next = v.prototypes[0].first_triangle;
        for (std::uint32_t k = 0; k < v.prototype_count; ++k) {
            const hdcx_prototype& p = v.prototypes[k];
            if (p.first_triangle != next || p.triangle_count == 0)
                throw Failure(HDCX_ERR_INVALID_SCENE, "prototypes must tile the triangles after the drawn ones");
            next += p.triangle_count;
        }
        if (next != v.triangle_count) throw Failure(HDCX_ERR_INVALID_SCENE, "prototypes must end at the last triangle");
    }
    if (v.instance_count) {
        if (!v.instances || !v.prototype_count) throw Failure(HDCX_ERR_INVALID_SCENE, "instances without prototypes");
        std::uint64_t flat = v.prototypes[0].first_triangle;
        for (std::uint32_t i = 0; i < v.instance_count; ++i) {
            const hdcx_instance& inst = v.instances[i];
            if (inst.prototype >= v.prototype_count) throw Failure(HDCX_ERR_INVALID_SCENE, "instance prototype out of range");
            if (inst.material != 0xffffffffU && inst.material >= v.material_count)
                throw Failure(HDCX_ERR_INVALID_SCENE, "instance material out of range");
            float toObject[12], normal[9];
            for (float x : inst.object_to_world)
                if (!std::isfinite(x)) throw Failure(HDCX_ERR_INVALID_SCENE, "non-finite instance transform");
            if (!InvertInstance(inst.object_to_world, toObject, normal))
                throw Failure(HDCX_ERR_INVALID_SCENE, "singular instance transform");
            flat += v.prototypes[inst.prototype].triangle_count;
        }
        if (flat >= 0xffffffffULL) throw Failure(HDCX_ERR_INVALID_SCENE, "more than 2^32 - 1 flattened triangles");
    }
    const std::uint32_t drawn = v.prototype_count ? v.prototypes[0].first_triangle : v.triangle_count;
    if (!v.emissive && v.light_triangle_count > drawn)
        throw Failure(HDCX_ERR_INVALID_SCENE, "light_triangle_count beyond the drawn triangles");
    const auto* ls = static_cast<const GpuLightRecord*>(v.lights);
    for (std::uint32_t l = 0; l < v.light_count; ++l) {
        ValidateTextureRef(ls[l].textureInfo.x, v.texture_count, "light");
        if (ls[l].textureInfo.w == kMissingTexture) continue;
        if (ls[l].textureInfo.z == 0U && ls[l].textureInfo.w == ls[l].textureInfo.x) {
            // Built on the device from the dome's own (plain) texture.
            if (ls[l].textureInfo.x & kUdimTextureFlag)
                throw Failure(HDCX_ERR_INVALID_SCENE, "dome importance from a UDIM set");
            continue;
        }
        if (ls[l].textureInfo.w >= v.texture_count || v.textures[ls[l].textureInfo.w].format != HDCX_TEXTURE_RGBA32F)
            throw Failure(HDCX_ERR_INVALID_SCENE, "light importance texture invalid");
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

// Instanced scene: a host SAH BVH per prototype and for the drawn
// triangles, and an instance tree over their world boxes; one node array
// (instance tree first, root 0) and one slot array.
void BuildTwoLevel(SceneStore* t, const hdcx_scene_view& v, hipStream_t st)
{
    const std::uint32_t drawn = v.prototypes[0].first_triangle;
    struct Blas {
        Bvh4 bvh;
        std::uint32_t nodeBase = 0, slotBase = 0;
        lbvh::Box4 box{};
    };
    auto rootBox = [](const Bvh4& b) {
        lbvh::Box4 box{};
        for (int k = 0; k < 3; ++k) {
            box.lo[k] = INFINITY;
            box.hi[k] = -INFINITY;
        }
        const Bvh4Node& r = b.nodes[0];
        for (int i = 0; i < 4; ++i) {
            if (r.child[i] == kEmptyChild) continue;
            box.lo[0] = std::min(box.lo[0], r.lox[i]);
            box.lo[1] = std::min(box.lo[1], r.loy[i]);
            box.lo[2] = std::min(box.lo[2], r.loz[i]);
            box.hi[0] = std::max(box.hi[0], r.hix[i]);
            box.hi[1] = std::max(box.hi[1], r.hiy[i]);
            box.hi[2] = std::max(box.hi[2], r.hiz[i]);
        }
        return box;
    };
    std::vector<Blas> blas(v.prototype_count + (drawn > 0 ? 1 : 0));
    for (std::uint32_t k = 0; k < v.prototype_count; ++k) {
        const hdcx_prototype& p = v.prototypes[k];
        Bvh4 b = BuildBvh4Host(v.positions, v.indices + static_cast<std::size_t>(p.first_triangle) * 3, p.triangle_count);
        ValidateBvh4(b, p.triangle_count);
        for (std::size_t slot = 0; slot < b.triangles.size() / 12; ++slot) {
            std::uint32_t id;
            std::memcpy(&id, &b.triangles[slot * 12 + 3], 4);
            id += p.first_triangle;
            std::memcpy(&b.triangles[slot * 12 + 3], &id, 4);
        }
        blas[k].box = rootBox(b);
        blas[k].bvh = std::move(b);
    }
    if (drawn > 0) {
        Bvh4 b = BuildBvh4Host(v.positions, v.indices, drawn);
        ValidateBvh4(b, drawn);
        blas.back().box = rootBox(b);
        blas.back().bvh = std::move(b);
    }
    // Instance boxes: the prototype box's corners through the transform
    // (FP64), widened by a bound on the FP32 vertex transform's rounding.
    std::vector<lbvh::Box4> items;
    std::vector<InstanceRecord> records(v.instance_count);
    std::vector<std::uint32_t> bases(v.instance_count);
    std::vector<float> pads(static_cast<std::size_t>(v.instance_count) * 12);
    std::uint32_t base = drawn;
    for (std::uint32_t i = 0; i < v.instance_count; ++i) {
        const hdcx_instance& in = v.instances[i];
        const lbvh::Box4& pb = blas[in.prototype].box;
        const float* m = in.object_to_world;
        InstanceRecord& r = records[i];
        std::memcpy(r.toWorld, m, sizeof(r.toWorld));
        if (!InvertInstance(m, r.toObject, r.normal)) throw Failure(HDCX_ERR_INVALID_SCENE, "singular instance transform");
        r.prototype = in.prototype;
        r.material = in.material;
        r.base = base;
        bases[i] = base;
        base += v.prototypes[in.prototype].triangle_count;
        double ext[3], objExt = 0.0;
        for (int k = 0; k < 3; ++k) {
            ext[k] = std::max(std::fabs(static_cast<double>(pb.lo[k])), std::fabs(static_cast<double>(pb.hi[k])));
            objExt = std::max(objExt, ext[k]);
        }
        lbvh::Box4 box{};
        double normM = 0.0, normInv = 0.0, tInf = 0.0;
        for (int row = 0; row < 3; ++row) {
            double lo = 0.0, hi = 0.0, err = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double a = m[4 * row + k];
                lo += std::min(a * pb.lo[k], a * pb.hi[k]);
                hi += std::max(a * pb.lo[k], a * pb.hi[k]);
                err += std::fabs(a) * ext[k];
            }
            const double tr = m[4 * row + 3];
            err = 8.0 * 5.9604644775390625e-8 * (err + std::fabs(tr));
            box.lo[row] = static_cast<float>(lo + tr - err);
            box.hi[row] = static_cast<float>(hi + tr + err);
            normM = std::max(normM, std::fabs(static_cast<double>(m[4 * row])) + std::fabs(static_cast<double>(m[4 * row + 1])) +
                                        std::fabs(static_cast<double>(m[4 * row + 2])));
            normInv = std::max(normInv, std::fabs(static_cast<double>(r.toObject[4 * row])) +
                                            std::fabs(static_cast<double>(r.toObject[4 * row + 1])) +
                                            std::fabs(static_cast<double>(r.toObject[4 * row + 2])));
            tInf = std::max(tInf, std::fabs(tr));
        }
        items.push_back(box);
        float* pad = &pads[12 * static_cast<std::size_t>(i)];
        pad[0] = static_cast<float>(normInv);
        pad[1] = static_cast<float>(normM);
        pad[2] = static_cast<float>(tInf);
        pad[3] = static_cast<float>(objExt);
        for (int k = 0; k < 3; ++k) {
            pad[4 + k] = box.lo[k];
            pad[8 + k] = box.hi[k];
        }
    }
    if (drawn > 0) items.push_back(blas.back().box);
    // World box of the scene: the instance boxes and the drawn triangles.
    for (int k = 0; k < 3; ++k) {
        float lo = items[0].lo[k], hi = items[0].hi[k];
        for (const lbvh::Box4& b : items) {
            lo = std::min(lo, b.lo[k]);
            hi = std::max(hi, b.hi[k]);
        }
        t->sceneLo[k] = lo;
        t->sceneInvExtent[k] = hi > lo ? 1.0F / (hi - lo) : 1.0F;
    }
    const Bvh4 tlas = BuildBvh4OverBoxes(items);
    // Layout: instance tree, then each BLAS; BLAS references relocated.
    std::uint32_t nodeCount = static_cast<std::uint32_t>(tlas.nodes.size()), slotCount = 0, blasStack = 0, blasDepth = 0;
    for (Blas& b : blas) {
        b.nodeBase = nodeCount;
        b.slotBase = slotCount;
        nodeCount += static_cast<std::uint32_t>(b.bvh.nodes.size());
        slotCount += static_cast<std::uint32_t>(b.bvh.triangles.size() / 12);
        blasStack = std::max(blasStack, b.bvh.maxStack);
        blasDepth = std::max(blasDepth, b.bvh.depth);
    }
    if (tlas.maxStack + 1 + blasStack > static_cast<std::uint32_t>(kMaxStack))
        throw Failure(HDCX_ERR_INVALID_SCENE, "instance tree plus BVH deeper than the traversal stack");
    std::vector<Bvh4Node> nodes(tlas.nodes);
    nodes.reserve(nodeCount);
    std::vector<float> slots;
    slots.reserve(static_cast<std::size_t>(slotCount) * 12);
    for (const Blas& b : blas) {
        for (Bvh4Node n : b.bvh.nodes) {
            for (std::uint32_t& c : n.child) {
                if (c == kEmptyChild) continue;
                if (c & kLeafFlag) c = EncodeLeaf(((c >> 3) & 0x0fffffffU) + b.slotBase, (c & 7U) + 1U);
                else c += b.nodeBase;
            }
            nodes.push_back(n);
        }
        slots.insert(slots.end(), b.bvh.triangles.begin(), b.bvh.triangles.end());
    }
    std::vector<PrototypeRecord> protos(v.prototype_count);
    for (std::uint32_t k = 0; k < v.prototype_count; ++k)
        protos[k] = PrototypeRecord{v.prototypes[k].first_triangle, v.prototypes[k].triangle_count, blas[k].nodeBase, 0U};
    t->nodes.Upload(nodes.data(), nodes.size() * sizeof(Bvh4Node), st);
    t->tris.Upload(slots.data(), slots.size() * sizeof(float), st);
    t->instanceRecords.Upload(records.data(), records.size() * sizeof(InstanceRecord), st);
    t->prototypeRecords.Upload(protos.data(), protos.size() * sizeof(PrototypeRecord), st);
    t->instanceBase.Upload(bases.data(), bases.size() * 4, st);
    t->instancePad.Upload(pads.data(), pads.size() * 4, st);
    Check(hipStreamSynchronize(st), "instance upload");
    t->refitPlan.order.Reset();
    t->refitPlan.begin.assign(1, 0U);
    t->bvhInfo = hdcx_bvh_info{};
    t->bvhInfo.node_count = nodeCount;
    t->bvhInfo.slot_count = slotCount;
    t->bvhInfo.max_stack = tlas.maxStack + 1 + blasStack;
    t->bvhInfo.depth = tlas.depth + blasDepth;
    t->bvhInfo.builder = HDCX_BVH_HOST_SAH;
    DeviceScene& d = t->scene;
    d.instanceCount = v.instance_count;
    d.drawnTriangles = drawn;
    d.tlasRoot = 0;
    d.drawnRoot = drawn > 0 ? blas.back().nodeBase : kEmptyChild;
    d.instances = t->instanceRecords.as<InstanceRecord>();
    d.prototypes = t->prototypeRecords.as<PrototypeRecord>();
    d.instanceBase = t->instanceBase.as<std::uint32_t>();
    d.instancePad = t->instancePad.as<float>();
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
    // Drawn triangles; prototype triangles only through instances.
    const std::uint32_t drawn = v.prototype_count ? v.prototypes[0].first_triangle : v.triangle_count;
    const bool instanced = v.instance_count > 0;
    if (!instanced && drawn == 0) {
        t->builtIndices.clear();
        t->builtVertexCount = 0;
        return;
    }
    const bool refit = !instanced && t->bvhOptions.refit != 0U && hadGeometry && t->builtVertexCount == v.vertex_count &&
                       t->builtIndices.size() == static_cast<std::size_t>(drawn) * 3 &&
                       std::memcmp(t->builtIndices.data(), v.indices, t->builtIndices.size() * 4) == 0;
    t->scene.instanceCount = 0;
    t->scene.drawnTriangles = drawn;
    t->scene.tlasRoot = 0;
    t->scene.drawnRoot = 0;
    t->positions.Upload(v.positions, static_cast<std::size_t>(v.vertex_count) * 16, st);
    t->indices.Upload(v.indices, static_cast<std::size_t>(v.triangle_count) * 12, st);
    const auto buildStart = std::chrono::steady_clock::now();
    if (refit) {
        RefitBvh4(t->positions.as<float4>(), t->indices.as<std::uint32_t>(), t->bvhInfo.slot_count,
                  t->nodes.as<Bvh4Node>(), t->tris.as<float4>(), t->refitPlan, st);
        ++t->bvhInfo.refits;
    } else if (instanced) {
        BuildTwoLevel(t, v, st);
        t->builtIndices.clear();
        t->builtVertexCount = 0;
    } else {
        hdcx_scene_view dv = v;
        dv.triangle_count = drawn;
        BuildSceneBvh(t, dv, lo, hi, st);
        t->builtIndices.assign(v.indices, v.indices + static_cast<std::size_t>(drawn) * 3);
        t->builtVertexCount = v.vertex_count;
    }
    t->bvhInfo.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - buildStart).count();
    t->triMaterial.Upload(v.triangle_materials, static_cast<std::size_t>(v.triangle_count) * 4, st);
    t->texcoords.Upload(v.texcoords, static_cast<std::size_t>(v.triangle_count) * 24, st);
    t->normals.Upload(v.normals, static_cast<std::size_t>(v.triangle_count) * 48, st);
    t->shading.Allocate(static_cast<std::size_t>(std::max<std::uint32_t>(v.triangle_count, 1)) * 128);
    Check(LaunchShadingRecords(t->positions.as<float4>(), t->indices.as<std::uint32_t>(), t->normals.as<float4>(),
                               t->texcoords.as<float2>(), t->triMaterial.as<std::uint32_t>(), v.triangle_count,
                               t->shading.as<float4>(), st),
          "shading records");
    const bool deviceEmissive = v.emissive == nullptr;
    if (!deviceEmissive) t->emissive.Upload(v.emissive, static_cast<std::size_t>(v.triangle_count) * 16, st);
    t->emissiveEntries = v.triangle_count;
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

    if (deviceEmissive) {
        // Emissive-triangle distribution on the device (:2944-2993).
        const std::uint32_t lightTriangles = v.light_triangle_count ? v.light_triangle_count : drawn;
        t->emissive.Allocate(static_cast<std::size_t>(v.triangle_count) * 16);
        t->materialPower.Allocate(static_cast<std::size_t>(v.material_count) * 4);
        if (v.material_emission_power) {
            t->materialPower.Upload(v.material_emission_power, static_cast<std::size_t>(v.material_count) * 4, st);
        } else {
            Check(LaunchMaterialPower(t->materials.as<GpuMaterialRecord>(), v.material_count, t->textures.as<TexDesc>(),
                                      v.texture_count, t->texels.as<unsigned char>(), t->materialPower.as<float>(), st),
                  "material power");
        }
        t->tableScratch.Reserve(EmissiveScratchDoubles(lightTriangles) * sizeof(double));
        Check(LaunchEmissiveTable(t->positions.as<float4>(), t->indices.as<std::uint32_t>(),
                                  t->triMaterial.as<std::uint32_t>(), lightTriangles, v.triangle_count,
                                  t->materialPower.as<float>(), v.material_count, t->emissive.as<float4>(),
                                  t->tableScratch.as<double>(), st),
              "emissive table");
    }

    // Dome importance tables per light: marginal CDF, conditional CDFs, pdf
    // plane and guide tables in one pool. Host-built importance textures
    // are re-laid out here; domes whose importance slot names their own
    // texture get theirs built on the device after the upload.
    const auto* ls = static_cast<const GpuLightRecord*>(v.lights);
    std::vector<ImportanceDesc> imp(lightRecords, ImportanceDesc{});
    std::vector<float> impPool;
    // Guide table of a CDF (2^shift + 1 uint32 entries appended to the pool),
    // or 0 when the CDF is not monotone (then the full binary search runs).
    auto appendGuide = [&impPool](const float* cdf, std::uint32_t n, std::uint32_t shift) -> bool {
        for (std::uint32_t i = 1; i < n; ++i)
            if (!(cdf[i] >= cdf[i - 1])) return false;
        const std::uint32_t buckets = 1U << shift;
        std::uint32_t i = 0;
        for (std::uint32_t g = 0; g <= buckets; ++g) {
            const float threshold = static_cast<float>(static_cast<double>(g) / buckets);  // exact
            while (i + 1U < n && cdf[i] < threshold) ++i;
            float word;
            std::memcpy(&word, &i, 4);
            impPool.push_back(word);
        }
        return true;
    };
    auto shiftFor = [](std::uint32_t n) {
        std::uint32_t k = 0;
        while ((1U << k) < n && k < 20U) ++k;
        return k;
    };
    std::vector<std::uint32_t> deviceDomes;
    for (std::uint32_t l = 0; l < v.light_count; ++l) {
        const std::uint32_t ti = ls[l].textureInfo.w;
        if (ls[l].textureInfo.z != 0U || ti == kMissingTexture) continue;
        if (ti == ls[l].textureInfo.x) {
            deviceDomes.push_back(l);
            continue;
        }
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
        d.rowShift = shiftFor(tv.height);
        d.colShift = shiftFor(tv.width);
        {
            const std::vector<float> marginal(impPool.begin() + static_cast<std::ptrdiff_t>(d.marginal),
                                              impPool.begin() + static_cast<std::ptrdiff_t>(d.marginal + tv.height));
            const std::size_t at = impPool.size();
            d.rowGuide = appendGuide(marginal.data(), tv.height, d.rowShift) ? at : 0;
            if (!d.rowGuide) impPool.resize(at);
        }
        {
            const std::size_t at = impPool.size();
            bool ok = true;
            std::vector<float> row(tv.width);
            for (std::uint32_t y = 0; y < tv.height && ok; ++y) {
                std::copy_n(impPool.begin() + static_cast<std::ptrdiff_t>(d.conditional + static_cast<std::size_t>(y) * tv.width),
                            tv.width, row.begin());
                ok = appendGuide(row.data(), tv.width, d.colShift);
            }
            d.colGuide = ok ? at : 0;
            if (!ok) impPool.resize(at);
        }
        imp[l] = d;
    }
    // Device-built tables follow the host part of the pool. Their CDFs are
    // monotone by construction (prefix sums of non-negative terms), so both
    // guide tables are always present.
    std::size_t poolFloats = std::max<std::size_t>(impPool.size(), 1);
    std::size_t scratchDoubles = 0;
    for (std::uint32_t l : deviceDomes) {
        const hdcx_texture_view& tv = v.textures[ls[l].textureInfo.x];
        ImportanceDesc d{};
        d.width = tv.width;
        d.height = tv.height;
        d.rowShift = shiftFor(tv.height);
        d.colShift = shiftFor(tv.width);
        const std::size_t texels = static_cast<std::size_t>(tv.width) * tv.height;
        d.marginal = poolFloats;
        d.conditional = d.marginal + tv.height;
        d.pdf = d.conditional + texels;
        d.rowGuide = d.pdf + texels;
        d.colGuide = d.rowGuide + (1U << d.rowShift) + 1U;
        poolFloats = d.colGuide + static_cast<std::size_t>(tv.height) * ((1U << d.colShift) + 1U);
        scratchDoubles = std::max(scratchDoubles, ImportanceScratchDoubles(tv.width, tv.height));
        imp[l] = d;
    }
    t->importance.Upload(imp.data(), imp.size() * sizeof(ImportanceDesc), st);
    t->importancePool.Allocate(poolFloats * sizeof(float));
    if (!impPool.empty())
        Check(hipMemcpyAsync(t->importancePool.as<float>(), impPool.data(), impPool.size() * sizeof(float),
                             hipMemcpyHostToDevice, st), "importance upload");
    if (!deviceDomes.empty()) {
        t->tableScratch.Reserve(scratchDoubles * sizeof(double));
        t->domeOk.Reserve(deviceDomes.size() * sizeof(std::uint32_t));
        for (std::size_t k = 0; k < deviceDomes.size(); ++k) {
            const std::uint32_t l = deviceDomes[k];
            const std::uint32_t ti = ls[l].textureInfo.x;
            Check(LaunchImportanceTables(descs[ti], t->texels.as<unsigned char>(), imp[l], t->importancePool.as<float>(),
                                         t->tableScratch.as<double>(), t->domeOk.as<std::uint32_t>() + k, st),
                  "dome importance tables");
        }
    }
    Check(hipStreamSynchronize(st), "scene upload");
    if (!deviceDomes.empty()) {
        // A dome without positive weight is sampled uniformly: its record's
        // importance slot becomes missing on the device.
        std::vector<std::uint32_t> ok(deviceDomes.size());
        Check(hipMemcpy(ok.data(), t->domeOk.as<std::uint32_t>(), ok.size() * 4, hipMemcpyDeviceToHost), "dome flags");
        for (std::size_t k = 0; k < deviceDomes.size(); ++k) {
            if (ok[k]) continue;
            const std::uint32_t l = deviceDomes[k];
            const std::uint32_t missing = kMissingTexture;
            Check(hipMemcpy(t->lights.as<unsigned char>() + static_cast<std::size_t>(l) * HDCX_LIGHT_BYTES +
                                offsetof(GpuLightRecord, textureInfo) + 3 * sizeof(std::uint32_t),
                            &missing, 4, hipMemcpyHostToDevice),
                  "dome record");
            imp[l] = ImportanceDesc{};
        }
    }
    t->importanceLayout.assign(imp.begin(), imp.begin() + v.light_count);
    float lastCdf = 0.0F;
    if (drawn > 0) {
        if (deviceEmissive)
            Check(hipMemcpy(&lastCdf, t->emissive.as<float>() + (static_cast<std::size_t>(drawn) - 1) * 4, 4,
                            hipMemcpyDeviceToHost), "emissive readback");
        else
            lastCdf = v.emissive[(static_cast<std::size_t>(drawn) - 1) * 4];
    }

    DeviceScene& s = t->scene;
    s.nodes = t->nodes.as<Bvh4Node>();
    s.tris = t->tris.as<float4>();
    s.positions = t->positions.as<float4>();
    s.indices = t->indices.as<std::uint32_t>();
    s.triMaterial = t->triMaterial.as<std::uint32_t>();
    s.texcoords = t->texcoords.as<float2>();
    s.normals = t->normals.as<float4>();
    s.emissive = t->emissive.as<float4>();
    s.shading = t->shading.as<float4>();
    s.materials = t->materials.as<GpuMaterialRecord>();
    s.lights = t->lights.as<GpuLightRecord>();
    s.textures = t->textures.as<TexDesc>();
    s.texels = t->texels.as<unsigned char>();
    s.importance = t->importance.as<ImportanceDesc>();
    s.importancePool = t->importancePool.as<float>();
    s.primitiveCount = drawn;  // emissive-triangle sampling covers the drawn triangles
    s.lightCount = v.light_count;
    s.opaque = v.opaque ? 1U : 0U;
    s.hasEmissive = lastCdf > 0.0F ? 1U : 0U;
    for (int k = 0; k < 3; ++k) {
        s.boundsLo[k] = t->sceneLo[k];
        s.boundsInvExtent[k] = t->sceneInvExtent[k];
    }
    s.hasSubsurface = 0U;
    s.simpleMaterials = 1U;
    const auto* mats = static_cast<const GpuMaterialRecord*>(v.materials);
    for (std::uint32_t m = 0; m < v.material_count; ++m) {
        const GpuMaterialRecord& r = mats[m];
        if (r.subsurfaceWeightScale.x > 0.0F || r.textureIndices2.x != kMissingTexture) s.hasSubsurface = 1U;
        // Every lobe the simple shading variant leaves out has weight 0
        // here (the weights are clamped to [0, 1]; NaN counts as 0).
        const bool plain = !(r.subsurfaceWeightScale.x > 0.0F) && r.textureIndices2.x == kMissingTexture &&
                           !(r.coatWeightRoughnessIor.x > 0.0F) && r.textureIndices3.z == kMissingTexture &&
                           !(r.coatColor.w > 0.0F) && r.textureIndices1.w == kMissingTexture &&
                           !(r.translucentColorWeight.w > 0.0F) && r.textureIndices5.x == kMissingTexture &&
                           !(r.transmissionOpacityIor.x > 0.0F);
        if (!plain) s.simpleMaterials = 0U;
    }
    s.textureCount = v.texture_count;
    s.nodeCount = t->bvhInfo.node_count;
    t->geometry = true;
}

void PathPool::EnsurePaths(std::uint32_t paths, std::uint32_t needs)
{
    if (!counters.bytes()) {
        counters.Allocate(kCounterSlots * kCounterStride * sizeof(std::uint32_t));
        totals.Allocate(64);
        Check(hipMemset(totals.as<void>(), 0, 64), "memset totals");
    }
    if (paths > capacity) {
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
        extendQueue0.Allocate(n * 4);
        extendQueue1.Allocate(n * 4);
        follow.Allocate(n * 4);
        shadowLeaders.Allocate(n * 4);
        shadowFollow.Allocate(n * 4);
        // The optional arrays are re-made at the new capacity when needed.
        walk.Reset();
        walkQueue.Reset();
        spread.Reset();
        pathScene.Reset();
        have = 0;
        capacity = paths;
    }
    const std::size_t n = capacity;
    if ((needs & kNeedWalks) && !(have & kNeedWalks)) {
        walk.Allocate(n * 48);
        walkQueue.Allocate(n * 4);
        have |= kNeedWalks;
    }
    if ((needs & kNeedSpread) && !(have & kNeedSpread)) {
        spread.Allocate(n * 8);
        have |= kNeedSpread;
    }
    if ((needs & kNeedSceneIndex) && !(have & kNeedSceneIndex)) {
        pathScene.Allocate(n * 4);
        have |= kNeedSceneIndex;
    }
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
    pb.walk = walk.as<float4>();
    pb.walkQueue = walkQueue.as<std::uint32_t>();
    pb.extendQueue[0] = extendQueue0.as<std::uint32_t>();
    pb.extendQueue[1] = extendQueue1.as<std::uint32_t>();
    pb.follow = follow.as<std::uint32_t>();
    pb.shadowLeaders = shadowLeaders.as<std::uint32_t>();
    pb.shadowFollow = shadowFollow.as<std::uint32_t>();
    pb.capacity = capacity;
    return pb;
}

} // namespace hdcx
