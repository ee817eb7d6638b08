// Packed record layouts shared by the host packer and the device kernels.
//
// GpuMaterialRecord and GpuLightRecord are byte-for-byte the reference's
// std430 structs (src/gpu/vulkan_path_tracer.cpp:42-66 / 1907-1933 and
// :71-86 / 1935-1952): the compiled MaterialX closure ABI and the light
// table are consumed unchanged, so the field names are kept.
#pragma once

#include <cstdint>

namespace hdcx {

struct alignas(16) F4 {
    float x, y, z, w;
};

struct alignas(16) U4 {
    std::uint32_t x, y, z, w;
};

struct GpuMaterialRecord {
    F4 baseColorMetalness;
    F4 emissionRoughness;
    F4 transmissionOpacityIor;
    F4 transmissionColorThinWalled;
    F4 subsurfaceWeightScale;
    F4 subsurfaceColor;
    F4 subsurfaceRadius;
    F4 specularWeightColor;
    F4 coatWeightRoughnessIor;
    F4 coatColor;
    F4 transmissionMedium;
    F4 transmissionScatter;
    F4 sheenColorMode;
    F4 diffuseProperties;
    F4 translucentColorWeight;
    F4 baseColorTextureScale;
    F4 baseColorTextureBias;
    U4 textureIndices0;
    U4 textureIndices1;
    U4 textureIndices2;
    U4 textureIndices3;
    U4 textureIndices4;
    U4 textureIndices5;
};
static_assert(sizeof(GpuMaterialRecord) == 368, "material record is 23 x 16 bytes");

struct GpuLightRecord {
    F4 colorIntensity;
    F4 controls;
    F4 position;
    F4 axisU;
    F4 axisV;
    F4 basisX;
    F4 basisY;
    F4 basisZ;
    F4 shaping;
    F4 focusTint;
    F4 shadow;
    F4 shadowColor;
    F4 geometry;
    U4 textureInfo;
};
static_assert(sizeof(GpuLightRecord) == 224, "light record is 14 x 16 bytes");

constexpr std::uint32_t kMissingTexture = 0xffffffffU;
constexpr std::uint32_t kUdimTextureFlag = 0x80000000U;
constexpr std::uint32_t kUdimTileMask = 0x7fffffU;
constexpr std::uint32_t kMaxTextures = 256U;
constexpr std::uint32_t kMaxLights = 64U;
constexpr std::uint32_t kMaxBounces = 12U;

} // namespace hdcx
