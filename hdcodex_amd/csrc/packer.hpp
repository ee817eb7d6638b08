// Host-side scene packer: SceneSnapshot -> the flattened buffers the device
// consumes (hdcx_scene_view).
#pragma once

#include "../../include/hdcx.h"
#include "../../include/hdcx_scene.hpp"
#include "records.h"

#include <cstdint>
#include <vector>

namespace hdcx {

struct PackedTexture {
    std::uint32_t width{0};
    std::uint32_t height{0};
    std::uint32_t format{HDCX_TEXTURE_RGBA8_UNORM};
    std::vector<std::uint8_t> rgba8;
    std::vector<float> rgba32f;
};

struct PackedScene {
    std::vector<float> positions;            // 4 per vertex
    std::vector<std::uint32_t> indices;      // 3 per triangle
    std::vector<std::uint32_t> triangleMaterials;
    std::vector<float> texcoords;            // 6 per triangle
    std::vector<float> normals;              // 12 per triangle
    std::vector<float> emissive;             // 4 per triangle; empty: built on the device
    std::vector<float> materialEmissionPower;  // 1 per material (device-built emissive table)
    std::uint32_t lightTriangleCount{0};     // mesh-light triangles [0, n); 0 = all drawn
    std::vector<GpuMaterialRecord> materials;
    std::vector<GpuLightRecord> lights;      // >= 1 (dummy when none authored)
    std::uint32_t lightCount{0};             // authored, visible, <= 64
    std::vector<PackedTexture> textures;
    std::vector<hdcx_texture_view> textureViews;
    bool opaque{true};
    std::vector<hdcx_prototype> prototypes;
    std::vector<hdcx_instance> instances;

    [[nodiscard]] bool HasGeometry() const noexcept
    {
        return !positions.empty() && !indices.empty();
    }
    // The view points into this object; refreshes textureViews.
    hdcx_scene_view View();
};

// The scene with every instance written out as ordinary triangles, in
// primitive-id order (what the reference would flatten): world vertices and
// normals from the same FP32 formulas the device uses on the fly.
PackedScene FlattenInstances(const hdcx_scene_view& view);

// Restates the reference's BuildScene host flattening and packing
// (src/gpu/vulkan_path_tracer.cpp:2604-2940). The light tables that the
// reference builds on the host next (emissive-triangle CDF, :2944-2993;
// lat-long dome importance, :1998-2066) are left to hdcx_set_scene, which
// builds them on the device: the view carries a NULL emissive table plus
// each material's emission power, and each importance-sampled dome names
// its own texture in its importance slot.
PackedScene PackScene(const SceneSnapshot& snapshot);

// src/gpu/vulkan_path_tracer.cpp:1968-1996
float TextureAverageLuminance(const SceneTexture& source);

} // namespace hdcx
