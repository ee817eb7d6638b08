// Scene light tables built on the device at SetScene: the emissive-triangle
// distribution (the reference's host loop, src/gpu/vulkan_path_tracer.cpp:
// 2944-2993) and the lat-long dome importance tables
// (BuildLatLongImportanceTexture, :1998-2066). Both are FP64 prefix sums
// over FP64 weights written as FP32, as the reference's host code does;
// the sums run in a fixed tree order (deterministic, within a few FP64 ulps
// of the reference's sequential order).
#pragma once

#include "dev_common.h"

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace hdcx {

// Doubles of scratch the emissive table needs for `triangles` triangles.
std::size_t EmissiveScratchDoubles(std::uint32_t triangles);

// out[p] = (cdf, probability / area, area, 0) for the first `drawn`
// triangles (weight = area x materialPower[triMaterial[p]], area from the
// FP64 cross product of the FP32 positions); triangles [drawn, all) are not
// sampled (cdf 1, pdf 0, area 0). Without any positive weight every entry is
// zero. materialPower has one entry per material (0 past `materials`).
hipError_t LaunchEmissiveTable(const float4* positions, const std::uint32_t* indices, const std::uint32_t* triMaterial,
                               std::uint32_t drawn, std::uint32_t all, const float* materialPower,
                               std::uint32_t materials, float4* out, double* scratch, hipStream_t st);

// Per-material emission power from the packed material records (for scene
// views that do not carry it): max(emissionWeight, 0) x the luminance
// 0.2126 r + 0.7152 g + 0.0722 b of max(emission, 0), or of the average
// texel of the emission texture (UDIM: the mean of the tiles' averages).
hipError_t LaunchMaterialPower(const GpuMaterialRecord* materials, std::uint32_t count, const TexDesc* textures,
                               std::uint32_t textureCount, const unsigned char* texels, float* power, hipStream_t st);

// The shading record of every triangle (8 float4 = one 128-byte line per
// triangle): r0..r2 the corner positions p0.xyz p1.xyz p2.xyz, r2.w..r4.y the
// corner normals n0.xyz n1.xyz n2.xyz, r4.z..r5.w the corner UVs, r6.x the
// material id bits; the rest zero. The values are the scene arrays' own,
// copied (positions through the vertex indices).
hipError_t LaunchShadingRecords(const float4* positions, const std::uint32_t* indices, const float4* normals,
                                const float2* texcoords, const std::uint32_t* triMaterial, std::uint32_t triangles,
                                float4* out, hipStream_t st);

// Doubles of scratch one dome's importance tables need (width x height texels).
std::size_t ImportanceScratchDoubles(std::uint32_t width, std::uint32_t height);

// Importance tables of one lat-long dome texture into the pool at d's
// offsets: marginal CDF of rows, conditional CDF per row, pdf per texel
// (luminance / total weight, per steradian), and the guide tables (entry g:
// first index whose CDF value is >= g / 2^shift). *ok = 1 when the total
// weight is positive and finite, else 0 (the dome is then sampled
// uniformly, as the reference does when it builds no importance texture).
hipError_t LaunchImportanceTables(const TexDesc& texture, const unsigned char* texels, const ImportanceDesc& d,
                                  float* pool, double* scratch, std::uint32_t* ok, hipStream_t st);

} // namespace hdcx
