// Scene packer. Each step cites the part of the reference BuildScene
// (src/gpu/vulkan_path_tracer.cpp) whose ordering and layout it keeps; the
// primitive ids the integrator reports are defined by this flattening.
#include "packer.hpp"
#include "instance_math.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>

namespace hdcx {
namespace {

float SrgbToLinear(float value)
{
    // IEC 61966-2-1 decode (:1961-1966).
    return value <= 0.04045F ? value / 12.92F
                             : std::pow((value + 0.055F) / 1.055F, 2.4F);
}

// Linear, finite, non-negative RGB of texel `pixel` (:1983-1992, :2023-2030).
void LinearTexel(const SceneTexture& t, std::size_t pixel, float rgb[3])
{
    for (std::size_t c = 0; c < 3; ++c) {
        float value = !t.rgbaFloat.empty()
            ? t.rgbaFloat[pixel * 4 + c]
            : static_cast<float>(t.rgba[pixel * 4 + c]) / 255.0F;
        if (t.rgbaFloat.empty() && t.srgb) value = SrgbToLinear(value);
        rgb[c] = std::isfinite(value) ? std::max(value, 0.0F) : 0.0F;
    }
}

bool TextureUsable(const SceneTexture& t)
{
    if (t.width == 0U || t.height == 0U) return false;
    const std::size_t texels = static_cast<std::size_t>(t.width) * t.height;
    if (!t.rgbaFloat.empty()) return t.rgbaFloat.size() >= texels * 4;
    return !t.rgba.empty() && t.rgba.size() >= texels * 4;
}

PackedTexture UploadFormat(const SceneTexture& t)
{
    // RGBA32F when float texels exist, else RGBA8 sRGB/UNORM (:2217-2221).
    PackedTexture p;
    p.width = t.width;
    p.height = t.height;
    if (!t.rgbaFloat.empty()) {
        p.format = HDCX_TEXTURE_RGBA32F;
        p.rgba32f = t.rgbaFloat;
        p.rgba32f.resize(static_cast<std::size_t>(t.width) * t.height * 4, 0.0F);
    } else {
        p.format = t.srgb ? HDCX_TEXTURE_RGBA8_SRGB : HDCX_TEXTURE_RGBA8_UNORM;
        p.rgba8 = t.rgba;
        p.rgba8.resize(static_cast<std::size_t>(t.width) * t.height * 4, 0);
    }
    return p;
}

F4 f4(float x, float y, float z, float w) { return F4{x, y, z, w}; }
F4 f4(const std::array<float, 3>& v, float w) { return F4{v[0], v[1], v[2], w}; }

GpuMaterialRecord DefaultMaterial(const std::array<float, 3>& color)
{
    // :2768-2800
    GpuMaterialRecord m{};
    m.baseColorMetalness = f4(color, 0.0F);
    m.emissionRoughness = f4(0.0F, 0.0F, 0.0F, 0.5F);
    m.transmissionOpacityIor = f4(0.0F, 1.0F, 1.5F, 0.0F);
    m.transmissionColorThinWalled = f4(1.0F, 1.0F, 1.0F, 0.0F);
    m.subsurfaceWeightScale = f4(0.0F, 1.0F, 0.0F, 0.0F);
    m.subsurfaceColor = f4(color, 0.0F);
    m.subsurfaceRadius = f4(1.0F, 0.2F, 0.1F, 0.0F);
    m.specularWeightColor = f4(1.0F, 1.0F, 1.0F, 1.0F);
    m.coatWeightRoughnessIor = f4(0.0F, 0.1F, 1.5F, 0.0F);
    m.coatColor = f4(1.0F, 1.0F, 1.0F, 0.0F);
    m.transmissionMedium = f4(0.0F, 0.0F, 0.0F, 20.0F);
    m.transmissionScatter = f4(0.0F, 0.0F, 0.0F, 0.0F);
    m.sheenColorMode = f4(1.0F, 1.0F, 1.0F, 0.0F);
    m.diffuseProperties = f4(0.0F, 0.0F, 1.0F, 0.0F);
    m.translucentColorWeight = f4(1.0F, 1.0F, 1.0F, 0.0F);
    m.baseColorTextureScale = f4(1.0F, 1.0F, 1.0F, -1.0F);
    m.baseColorTextureBias = f4(0.0F, 0.0F, 0.0F, 0.0F);
    const U4 none{kMissingTexture, kMissingTexture, kMissingTexture, kMissingTexture};
    m.textureIndices0 = m.textureIndices1 = m.textureIndices2 = none;
    m.textureIndices3 = m.textureIndices4 = m.textureIndices5 = none;
    return m;
}

} // namespace

float TextureAverageLuminance(const SceneTexture& source)
{
    if (!TextureUsable(source)) return 0.0F;
    const std::size_t texels = static_cast<std::size_t>(source.width) * source.height;
    double total = 0.0;
    for (std::size_t p = 0; p < texels; ++p) {
        float rgb[3];
        LinearTexel(source, p, rgb);
        total += 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2];
    }
    return static_cast<float>(total / static_cast<double>(texels));
}

PackedScene FlattenInstances(const hdcx_scene_view& v)
{
    PackedScene out;
    const std::uint32_t drawn = v.prototype_count ? v.prototypes[0].first_triangle : v.triangle_count;
    out.positions.assign(v.positions, v.positions + static_cast<std::size_t>(v.vertex_count) * 4);
    out.indices.assign(v.indices, v.indices + static_cast<std::size_t>(drawn) * 3);
    out.triangleMaterials.assign(v.triangle_materials, v.triangle_materials + drawn);
    out.texcoords.assign(v.texcoords, v.texcoords + static_cast<std::size_t>(drawn) * 6);
    out.normals.assign(v.normals, v.normals + static_cast<std::size_t>(drawn) * 12);
    if (v.emissive) {
        out.emissive.assign(v.emissive, v.emissive + static_cast<std::size_t>(drawn) * 4);
    } else {
        // The device builds the table; the instances' copies stay out of it.
        out.lightTriangleCount = v.light_triangle_count ? v.light_triangle_count : drawn;
        if (v.material_emission_power)
            out.materialEmissionPower.assign(v.material_emission_power, v.material_emission_power + v.material_count);
    }
    for (std::uint32_t i = 0; i < v.instance_count; ++i) {
        const hdcx_instance& inst = v.instances[i];
        const hdcx_prototype& proto = v.prototypes[inst.prototype];
        float normal[9], toObject[12];
        if (!InvertInstance(inst.object_to_world, toObject, normal))
            throw std::invalid_argument("singular instance transform");
        for (std::uint32_t k = 0; k < proto.triangle_count; ++k) {
            const std::uint32_t t = proto.first_triangle + k;
            const auto base = static_cast<std::uint32_t>(out.positions.size() / 4);
            for (int c = 0; c < 3; ++c) {
                const float* p = v.positions + static_cast<std::size_t>(v.indices[3 * t + c]) * 4;
                float w[3];
                TransformPoint(inst.object_to_world, p[0], p[1], p[2], w);
                out.positions.insert(out.positions.end(), {w[0], w[1], w[2], 1.0F});
                out.indices.push_back(base + static_cast<std::uint32_t>(c));
                const float* n = v.normals + static_cast<std::size_t>(t) * 12 + 4 * c;
                float wn[3];
                TransformNormal(normal, n[0], n[1], n[2], wn);
                out.normals.insert(out.normals.end(), {wn[0], wn[1], wn[2], 0.0F});
            }
            out.triangleMaterials.push_back(inst.material == 0xffffffffU ? v.triangle_materials[t] : inst.material);
            out.texcoords.insert(out.texcoords.end(), v.texcoords + static_cast<std::size_t>(t) * 6,
                                 v.texcoords + static_cast<std::size_t>(t) * 6 + 6);
            // Never sampled as a light (as on the device): the final CDF value, zero pdf.
            if (v.emissive)
                out.emissive.insert(out.emissive.end(), {v.emissive[static_cast<std::size_t>(t) * 4], 0.0F, 0.0F, 0.0F});
        }
    }
    const auto* mats = static_cast<const GpuMaterialRecord*>(v.materials);
    out.materials.assign(mats, mats + v.material_count);
    const auto* lights = static_cast<const GpuLightRecord*>(v.lights);
    out.lights.assign(lights, lights + std::max<std::uint32_t>(v.light_count, 1U));
    out.lightCount = v.light_count;
    for (std::uint32_t i = 0; i < v.texture_count; ++i) {
        const hdcx_texture_view& tv = v.textures[i];
        PackedTexture t;
        t.width = tv.width;
        t.height = tv.height;
        t.format = tv.format;
        const std::size_t n = static_cast<std::size_t>(tv.width) * tv.height * 4;
        if (tv.format == HDCX_TEXTURE_RGBA32F) {
            const auto* f = static_cast<const float*>(tv.texels);
            t.rgba32f.assign(f, f + n);
        } else {
            const auto* b = static_cast<const std::uint8_t*>(tv.texels);
            t.rgba8.assign(b, b + n);
        }
        out.textures.push_back(std::move(t));
    }
    out.opaque = v.opaque != 0U;
    return out;
}

hdcx_scene_view PackedScene::View()
{
    textureViews.resize(textures.size());
    for (std::size_t i = 0; i < textures.size(); ++i) {
        const PackedTexture& t = textures[i];
        textureViews[i] = hdcx_texture_view{
            t.width, t.height, t.format, 0U,
            t.format == HDCX_TEXTURE_RGBA32F
                ? static_cast<const void*>(t.rgba32f.data())
                : static_cast<const void*>(t.rgba8.data())};
    }
    hdcx_scene_view v{};
    v.vertex_count = static_cast<std::uint32_t>(positions.size() / 4);
    v.triangle_count = static_cast<std::uint32_t>(triangleMaterials.size());
    v.positions = positions.data();
    v.indices = indices.data();
    v.triangle_materials = triangleMaterials.data();
    v.texcoords = texcoords.data();
    v.normals = normals.data();
    v.emissive = emissive.empty() ? nullptr : emissive.data();
    v.material_emission_power = materialEmissionPower.empty() ? nullptr : materialEmissionPower.data();
    v.light_triangle_count = lightTriangleCount;
    v.material_count = static_cast<std::uint32_t>(materials.size());
    v.light_count = lightCount;
    v.materials = materials.data();
    v.lights = lights.data();
    v.texture_count = static_cast<std::uint32_t>(textures.size());
    v.opaque = opaque ? 1U : 0U;
    v.textures = textureViews.data();
    v.prototype_count = static_cast<std::uint32_t>(prototypes.size());
    v.prototypes = prototypes.data();
    v.instance_count = static_cast<std::uint32_t>(instances.size());
    v.instances = instances.data();
    return v;
}

PackedScene PackScene(const SceneSnapshot& snapshot)
{
    PackedScene out;
    // Opaque iff every material is opaque without an opacity texture (:2609-2612).
    out.opaque = std::all_of(snapshot.materials.begin(), snapshot.materials.end(),
        [](const SceneMaterial& m) { return m.opacity >= 0.999F && m.opacityTexture.empty(); });

    // Textures: plain textures in snapshot order (<= 256), then UDIM sets in
    // set-id order with tiles sorted by tile number (:2627-2654).
    std::map<std::string, std::uint32_t, std::less<>> textureIndex;
    std::map<std::string, const SceneTexture*, std::less<>> plainSource;
    std::map<std::string, std::vector<const SceneTexture*>, std::less<>> udimSets;
    for (const SceneTexture& t : snapshot.textures) {
        if (!t.udimSetId.empty() && t.udimTile != 0U) {
            udimSets[t.udimSetId].push_back(&t);
            continue;
        }
        if (out.textures.size() >= kMaxTextures) continue;
        plainSource[t.id] = &t;
        textureIndex[t.id] = static_cast<std::uint32_t>(out.textures.size());
        out.textures.push_back(UploadFormat(t));
    }
    for (auto& [setId, tiles] : udimSets) {
        std::sort(tiles.begin(), tiles.end(), [](const SceneTexture* a, const SceneTexture* b) {
            return a->udimTile < b->udimTile;
        });
        const std::uint32_t base = static_cast<std::uint32_t>(out.textures.size());
        std::uint32_t present = 0U;
        for (const SceneTexture* tile : tiles) {
            if (out.textures.size() >= kMaxTextures) break;
            if (tile->udimTile < 1001U || tile->udimTile > 1023U) continue;
            present |= 1U << (tile->udimTile - 1001U);
            out.textures.push_back(UploadFormat(*tile));
        }
        if (present != 0U) {
            textureIndex[setId] = kUdimTextureFlag | ((present & kUdimTileMask) << 8U) | base;
        }
    }
    const auto indexOf = [&](const std::string& id) {
        const auto found = textureIndex.find(id);
        return found == textureIndex.end() ? kMissingTexture : found->second;
    };
    std::map<std::string, float, std::less<>> luminanceCache;
    const auto averageLuminance = [&](const std::string& id) {
        if (const auto c = luminanceCache.find(id); c != luminanceCache.end()) return c->second;
        double total = 0.0;
        std::size_t count = 0;
        if (const auto s = plainSource.find(id); s != plainSource.end()) {
            total = TextureAverageLuminance(*s->second);
            count = 1;
        } else if (const auto set = udimSets.find(id); set != udimSets.end()) {
            for (const SceneTexture* tile : set->second) {
                total += TextureAverageLuminance(*tile);
                ++count;
            }
        }
        const float value = count > 0 ? static_cast<float>(total / static_cast<double>(count)) : 0.0F;
        luminanceCache.emplace(id, value);
        return value;
    };
    // A lat-long (or automatic) dome with a usable plain texture is
    // importance sampled. The reference appends one importance texture per
    // distinct dome texture, in the order the domes first reference them,
    // while fewer than 256 textures exist (:2681-2702); those slots are
    // counted here, and the dome's importance slot names its own texture:
    // the tables are built on the device at hdcx_set_scene.
    std::map<std::string, std::uint32_t, std::less<>> importanceSlot;
    std::size_t textureSlots = out.textures.size();
    const auto domeImportance = [&](const SceneLight& light) {
        if (light.type != SceneLightType::Dome ||
            (light.textureFormat != DomeTextureFormat::Automatic &&
             light.textureFormat != DomeTextureFormat::LatLong) ||
            indexOf(light.texture) == kMissingTexture) {
            return kMissingTexture;
        }
        if (const auto c = importanceSlot.find(light.texture); c != importanceSlot.end()) return c->second;
        const auto s = plainSource.find(light.texture);
        if (s == plainSource.end() || textureSlots >= kMaxTextures || !TextureUsable(*s->second)) return kMissingTexture;
        ++textureSlots;
        const std::uint32_t index = indexOf(light.texture);
        importanceSlot.emplace(light.texture, index);
        return index;
    };

    // Lights: visible lights in snapshot order, at most 64 (:2704-2766).
    for (const SceneLight& light : snapshot.lights) {
        if (!light.visible || out.lights.size() >= kMaxLights) continue;
        float radianceScale = light.intensity * std::exp2(light.exposure);
        if (light.type != SceneLightType::Dome && light.type != SceneLightType::Distant &&
            light.normalize) {
            radianceScale /= std::max(light.area, 1e-8F);
        }
        const std::array<float, 3> emitted{light.color[0] * light.temperatureColor[0],
                                           light.color[1] * light.temperatureColor[1],
                                           light.color[2] * light.temperatureColor[2]};
        float emittedLuminance = 0.2126F * std::max(emitted[0], 0.0F) +
                                 0.7152F * std::max(emitted[1], 0.0F) +
                                 0.0722F * std::max(emitted[2], 0.0F);
        if (!light.texture.empty() &&
            (light.type == SceneLightType::Dome || light.type == SceneLightType::Rect)) {
            emittedLuminance *= averageLuminance(light.texture);
        }
        float measure = 1.0F;
        if (light.type == SceneLightType::Dome) measure = 12.5663706144F;
        else if (light.type != SceneLightType::Distant) measure = std::max(light.area, 0.0F);
        float power = std::max(emittedLuminance * std::max(radianceScale, 0.0F) * measure, 0.0F);
        if (!std::isfinite(power)) power = 0.0F;
        GpuLightRecord r{};
        r.colorIntensity = f4(emitted, radianceScale);
        r.controls = f4(light.diffuse, light.specular, static_cast<float>(light.type),
                        std::max(light.area, 0.0F));
        r.position = f4(light.position, 0.0F);
        r.axisU = f4(light.axisU, 0.0F);
        r.axisV = f4(light.axisV, 0.0F);
        r.basisX = f4(light.basisX, 0.0F);
        r.basisY = f4(light.basisY, 0.0F);
        r.basisZ = f4(light.basisZ, 0.0F);
        r.shaping = f4(light.shapingFocus, light.shapingConeAngle, light.shapingConeSoftness, 0.0F);
        r.focusTint = f4(light.shapingFocusTint, 0.0F);
        r.shadow = f4(light.shadowEnable ? 1.0F : 0.0F, light.shadowDistance,
                      light.shadowFalloff, light.shadowFalloffGamma);
        r.shadowColor = f4(light.shadowColor, 0.0F);
        r.geometry = f4(light.radius, light.length, light.angle, power);
        r.textureInfo = U4{indexOf(light.texture), static_cast<std::uint32_t>(light.textureFormat),
                           static_cast<std::uint32_t>(light.type), domeImportance(light)};
        out.lights.push_back(r);
    }
    out.lightCount = static_cast<std::uint32_t>(out.lights.size());
    if (out.lights.empty()) {
        GpuLightRecord dummy{};
        dummy.shadow = f4(1.0F, -1.0F, -1.0F, 1.0F);
        out.lights.push_back(dummy);
    }

    // Materials: index 0 is the default grey, then snapshot order (:2801-2881).
    std::vector<float> emissionPower;
    out.materials.push_back(DefaultMaterial({0.8F, 0.8F, 0.8F}));
    emissionPower.push_back(0.0F);
    std::map<std::string, std::uint32_t, std::less<>> materialIndex;
    for (const SceneMaterial& m : snapshot.materials) {
        materialIndex[m.id] = static_cast<std::uint32_t>(out.materials.size());
        GpuMaterialRecord r{};
        r.baseColorMetalness = f4(m.baseColor, m.metalness);
        r.emissionRoughness = f4(m.emission, m.roughness);
        r.transmissionOpacityIor = f4(m.transmission, m.opacity, m.indexOfRefraction, m.emissionWeight);
        r.transmissionColorThinWalled = f4(m.transmissionColor, m.thinWalled ? 1.0F : 0.0F);
        r.subsurfaceWeightScale = f4(m.subsurface, m.subsurfaceScale,
                                     m.subsurfaceScatterAnisotropy, m.normalTextureScale);
        r.subsurfaceColor = f4(m.subsurfaceColor, m.roughnessV);
        r.subsurfaceRadius = f4(m.subsurfaceRadius, m.coatRoughnessV);
        r.specularWeightColor = f4(m.specularWeight, m.specularColor[0], m.specularColor[1],
                                   m.specularColor[2]);
        r.coatWeightRoughnessIor = f4(m.coat, m.coatRoughness, m.coatIndexOfRefraction,
                                      m.normalTextureFlipY ? 1.0F : 0.0F);
        r.coatColor = f4(m.coatColor, m.sheen);
        r.transmissionMedium = f4(m.transmissionDepth, m.transmissionScatterAnisotropy,
                                  m.transmissionDispersionScale, m.transmissionDispersionAbbeNumber);
        r.transmissionScatter = f4(m.transmissionScatter, m.sheenRoughness);
        r.sheenColorMode = f4(m.sheenColor, static_cast<float>(m.sheenMode));
        r.diffuseProperties = f4(m.diffuseRoughness, static_cast<float>(m.diffuseModel),
                                 m.diffuseWeight, 0.0F);
        r.translucentColorWeight = f4(m.translucentColor, m.translucentWeight);
        r.baseColorTextureScale = f4(m.baseColorTextureScale,
                                     static_cast<float>(m.baseColorTextureChannel));
        r.baseColorTextureBias = f4(m.baseColorTextureBias, 0.0F);
        r.textureIndices0 = U4{indexOf(m.baseColorTexture), indexOf(m.metalnessTexture),
                               indexOf(m.roughnessTexture), indexOf(m.emissionTexture)};
        r.textureIndices1 = U4{indexOf(m.opacityTexture), indexOf(m.normalTexture),
                               indexOf(m.transmissionTexture), indexOf(m.sheenTexture)};
        r.textureIndices2 = U4{indexOf(m.subsurfaceTexture), indexOf(m.subsurfaceColorTexture),
                               indexOf(m.subsurfaceRadiusTexture), indexOf(m.sheenColorTexture)};
        r.textureIndices3 = U4{indexOf(m.specularTexture), indexOf(m.specularColorTexture),
                               indexOf(m.coatTexture), indexOf(m.coatColorTexture)};
        r.textureIndices4 = U4{indexOf(m.coatRoughnessTexture), indexOf(m.sheenRoughnessTexture),
                               indexOf(m.diffuseRoughnessTexture), indexOf(m.diffuseWeightTexture)};
        r.textureIndices5 = U4{indexOf(m.translucentWeightTexture),
                               indexOf(m.translucentColorTexture), kMissingTexture, kMissingTexture};
        out.materials.push_back(r);
        const float luminance = m.emissionTexture.empty()
            ? 0.2126F * std::max(m.emission[0], 0.0F) + 0.7152F * std::max(m.emission[1], 0.0F) +
                  0.0722F * std::max(m.emission[2], 0.0F)
            : averageLuminance(m.emissionTexture);
        emissionPower.push_back(std::max(m.emissionWeight, 0.0F) * luminance);
    }

    // Meshes flattened in snapshot order; triangles with an out-of-range
    // index are dropped; meshes without a material binding share one
    // default material per clamped display color (:2882-2940).
    std::map<std::array<float, 3>, std::uint32_t> displayColorMaterial;
    std::size_t vertexTotal = 0;
    for (const SceneMesh& mesh : snapshot.meshes) vertexTotal += mesh.positions.size() / 3;
    out.positions.reserve(vertexTotal * 4);
    // Drawn meshes first, then prototype meshes (each a range of triangles
    // after the drawn ones), both in snapshot order.
    std::vector<const SceneMesh*> order;
    for (const SceneMesh& mesh : snapshot.meshes)
        if (!mesh.prototype) order.push_back(&mesh);
    const std::size_t drawnMeshes = order.size();
    for (const SceneMesh& mesh : snapshot.meshes)
        if (mesh.prototype) order.push_back(&mesh);
    std::map<std::string, std::uint32_t> prototypeIndex;
    for (std::size_t oi = 0; oi < order.size(); ++oi) {
        const SceneMesh& mesh = *order[oi];
        const auto firstTriangle = static_cast<std::uint32_t>(out.triangleMaterials.size());
        const std::uint32_t base = static_cast<std::uint32_t>(out.positions.size() / 4);
        for (std::size_t i = 0; i + 2 < mesh.positions.size(); i += 3) {
            out.positions.insert(out.positions.end(),
                                 {mesh.positions[i], mesh.positions[i + 1], mesh.positions[i + 2], 1.0F});
        }
        const std::uint32_t meshVertices = static_cast<std::uint32_t>(out.positions.size() / 4) - base;
        const auto bound = materialIndex.find(mesh.materialId);
        std::uint32_t meshMaterial = bound == materialIndex.end() ? 0U : bound->second;
        if (mesh.materialId.empty()) {
            std::array<float, 3> color = mesh.displayColor;
            for (float& c : color) c = std::isfinite(c) ? std::clamp(c, 0.0F, 1.0F) : 0.5F;
            const auto found = displayColorMaterial.find(color);
            if (found != displayColorMaterial.end()) {
                meshMaterial = found->second;
            } else {
                meshMaterial = static_cast<std::uint32_t>(out.materials.size());
                displayColorMaterial.emplace(color, meshMaterial);
                out.materials.push_back(DefaultMaterial(color));
                emissionPower.push_back(0.0F);
            }
        }
        for (std::size_t tri = 0; tri + 2 < mesh.indices.size(); tri += 3) {
            const std::uint32_t i0 = mesh.indices[tri];
            const std::uint32_t i1 = mesh.indices[tri + 1];
            const std::uint32_t i2 = mesh.indices[tri + 2];
            if (i0 >= meshVertices || i1 >= meshVertices || i2 >= meshVertices) continue;
            out.indices.insert(out.indices.end(), {base + i0, base + i1, base + i2});
            const std::size_t triangleIndex = tri / 3;
            std::uint32_t material = meshMaterial;
            if (triangleIndex < mesh.triangleMaterialIds.size() &&
                mesh.triangleMaterialIds[triangleIndex] != mesh.materialId) {
                const auto f = materialIndex.find(mesh.triangleMaterialIds[triangleIndex]);
                material = f == materialIndex.end() ? 0U : f->second;
            }
            out.triangleMaterials.push_back(material);
            for (std::size_t corner = 0; corner < 3; ++corner) {
                const std::size_t uv = (tri + corner) * 2;
                if (uv + 1 < mesh.texcoords.size()) {
                    out.texcoords.insert(out.texcoords.end(), {mesh.texcoords[uv], mesh.texcoords[uv + 1]});
                } else {
                    out.texcoords.insert(out.texcoords.end(), {0.0F, 0.0F});
                }
                const std::size_t n = (tri + corner) * 3;
                if (n + 2 < mesh.normals.size()) {
                    out.normals.insert(out.normals.end(),
                                       {mesh.normals[n], mesh.normals[n + 1], mesh.normals[n + 2], 0.0F});
                } else {
                    out.normals.insert(out.normals.end(), {0.0F, 0.0F, 0.0F, 0.0F});
                }
            }
        }
        if (oi >= drawnMeshes) {
            const auto count = static_cast<std::uint32_t>(out.triangleMaterials.size()) - firstTriangle;
            if (count > 0) {
                prototypeIndex.emplace(mesh.id, static_cast<std::uint32_t>(out.prototypes.size()));
                out.prototypes.push_back(hdcx_prototype{firstTriangle, count});
            }
        }
    }
    // Instances in snapshot order; unknown prototypes are dropped, unknown
    // materials fall back to the prototype's own.
    for (const SceneInstance& inst : snapshot.instances) {
        const auto proto = prototypeIndex.find(inst.prototypeMeshId);
        if (proto == prototypeIndex.end()) continue;
        bool finite = true;
        for (float x : inst.objectToWorld) finite = finite && std::isfinite(x);
        if (!finite) continue;
        hdcx_instance r{};
        std::copy(inst.objectToWorld.begin(), inst.objectToWorld.end(), r.object_to_world);
        r.prototype = proto->second;
        const auto mat = inst.materialId.empty() ? materialIndex.end() : materialIndex.find(inst.materialId);
        r.material = mat == materialIndex.end() ? 0xffffffffU : mat->second;
        out.instances.push_back(r);
    }
    if (out.instances.empty()) {
        // Prototypes that no instance draws are not part of the scene.
        const std::uint32_t drawn = out.prototypes.empty() ? static_cast<std::uint32_t>(out.triangleMaterials.size())
                                                          : out.prototypes.front().first_triangle;
        out.triangleMaterials.resize(drawn);
        out.indices.resize(static_cast<std::size_t>(drawn) * 3);
        out.texcoords.resize(static_cast<std::size_t>(drawn) * 6);
        out.normals.resize(static_cast<std::size_t>(drawn) * 12);
        out.prototypes.clear();
    }
    if (out.positions.empty() || out.indices.empty()) {
        out.positions.clear();
        out.indices.clear();
        out.triangleMaterials.clear();
        out.texcoords.clear();
        out.normals.clear();
        out.prototypes.clear();
        out.instances.clear();
        return out;
    }

    // The emissive-triangle table (:2944-2993) is built on the device from
    // these powers; prototype triangles are never sampled as lights.
    out.materialEmissionPower = std::move(emissionPower);
    return out;
}

} // namespace hdcx

// ---- C++ wrapper support (include/hdcx_path_tracer.hpp) ----
#include "../../include/hdcx_path_tracer.hpp"

namespace hdcx {

PackedSceneHandle::PackedSceneHandle(const SceneSnapshot& snapshot) : packed_(new PackedScene(PackScene(snapshot)))
{
    if (packed_->HasGeometry()) view_ = packed_->View();
}

PackedSceneHandle::~PackedSceneHandle() { delete packed_; }

} // namespace hdcx
