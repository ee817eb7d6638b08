// Device-side types and helpers of the wavefront integrator (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/hdcx.h"
#include "bvh.hpp"
#include "records.h"
#include "instance_math.h"

namespace hdcx {

// ---------------------------------------------------------------------------
// Scene as seen by the kernels. All pointers are device memory owned by the
// tracer; the struct itself is passed by value as a kernel argument.
// ---------------------------------------------------------------------------
struct TexDesc {
    std::uint64_t offset;  // bytes into the texel pool
    std::uint32_t width, height, format, pad;
};

// Dome importance tables in a compact layout: the marginal CDF of rows
// (height floats), and per texel the conditional CDF and the solid-angle
// pdf. Values are those of the packed importance texture (r, g, b).
struct ImportanceDesc {
    std::uint64_t marginal, conditional, pdf;  // float offsets into the pool
    std::uint32_t width, height;
    // Guide tables (uint32 words in the pool), 0 when a CDF is not monotone:
    // rowGuide has 2^rowShift + 1 entries, colGuide 2^colShift + 1 per row;
    // entry g is the first index whose CDF value is >= g / 2^shift.
    std::uint64_t rowGuide, colGuide;
    std::uint32_t rowShift, colShift;
};

struct DeviceScene {
    const Bvh4Node* nodes;
    const float4* tris;              // 3 per leaf slot
    const float4* positions;
    const std::uint32_t* indices;
    const std::uint32_t* triMaterial;
    const float2* texcoords;         // 3 per triangle
    const float4* normals;           // 3 per triangle
    const float4* emissive;          // 1 per triangle
    // Per triangle, what shading reads of it in one 128-byte line (8 float4,
    // see ShadingRecord): corner positions, corner normals, UVs, material.
    const float4* shading;
    const GpuMaterialRecord* materials;
    const GpuLightRecord* lights;
    const TexDesc* textures;
    const unsigned char* texels;
    const ImportanceDesc* importance;  // indexed by light
    const float* importancePool;
    std::uint32_t primitiveCount;
    std::uint32_t lightCount;
    std::uint32_t opaque;
    std::uint32_t zUp;
    std::uint32_t hasEmissive;
    std::uint32_t textureCount;
    std::uint32_t nodeCount;
    std::uint32_t hasSubsurface;       // some material can take the subsurface random walk
    std::uint32_t simpleMaterials;     // no material has coat, sheen, transmission, subsurface or translucency
    float boundsLo[3];                 // world box of the scene (radiance-cache position normalisation)
    float boundsInvExtent[3];
    // Two-level scenes (instancing); instanceCount == 0 for one level.
    std::uint32_t instanceCount;
    std::uint32_t drawnTriangles;      // primitive ids below this are drawn triangles
    std::uint32_t tlasRoot;            // instance tree root (node 0)
    std::uint32_t drawnRoot;           // BLAS of the drawn triangles, kEmptyChild when none
    const InstanceRecord* instances;
    const PrototypeRecord* prototypes;
    const std::uint32_t* instanceBase; // flattened id of each instance's first triangle
    const float* instancePad;          // per instance: 12 floats: padding bound terms, world box lo, hi
};

// ---------------------------------------------------------------------------
// Path state of one chunk, structure of arrays indexed by path id.
// ---------------------------------------------------------------------------
enum PathFlags : std::uint32_t {
    kEnvOnMiss = 1U << 0,
    kPrevDelta = 1U << 1,
    kInMedium = 1U << 2,
};

constexpr int kShadowSlots = 3;
constexpr std::uint32_t kShadowMeshLight = 0xfffffffeU;  // boolean visibility, no shadow controls

constexpr std::uint32_t kCounterStride = 32;  // 32 x 4 bytes: one queue size per 128-byte line
constexpr std::uint32_t kHitCounter = 7;      // counter slot of the hit queue
constexpr std::uint32_t kCounterSlots = 8;

struct PathBuffers {
    float4* rayO;            // origin xyz
    float4* rayD;            // direction xyz
    float4* hit;             // t, primitive bits, u, v
    std::uint32_t* rng;
    float* wavelength;
    float* throughput;
    float* radiance;
    std::uint32_t* flags;
    float2* mis;             // previousBsdfPdf, previousEnvironmentLightPdf
    float4* medium;          // absorption, scattering, anisotropy
    float4* shadowO;         // [slot * capacity + path]: origin xyz, tmax
    float4* shadowD;         // direction xyz, unoccluded contribution
    std::uint32_t* shadowLight;  // light index, kShadowMeshLight, or ~0 = empty
    std::uint32_t* queue[2];
    std::uint32_t* shadowQueue;
    std::uint32_t* extendQueue[2];  // lane sharing: the surviving paths that trace their ray
    std::uint32_t* follow;   // lane sharing: bit j = path pid + 1 + j takes this path's hit
    std::uint32_t* shadowLeaders;  // shadow sharing: the queued paths that trace their shadow rays
    std::uint32_t* shadowFollow;   // shadow sharing: bit j = path pid + 1 + j shares them
    // Queue sizes, one per 128-byte line (index k at counters[k * kCounterStride],
    // so that the per-wave atomics on different queues do not share a line):
    // [0],[1] path queue sizes; [2] shadow queue size; [3] walk queue size;
    // [4],[5] extendQueue sizes; [6] shadowLeaders size; [7] hit queue size
    // (the hit queue of bounce `in` reuses extendQueue[in] once extend has
    // read it).
    std::uint32_t* counters;
    unsigned long long* totals;  // [0] paths extended, [1] shadow-queued paths, [2] shadow rays traced,
                                 // [3] extend rays traced (lanes of a pixel sample trace an equal ray once),
                                 // [4] paths shaded at a surface hit
    float* lambda0;          // wavelength of lane 0 of the path's hero triple
    float2* spread;          // radiance cache: primary footprint a0, accumulated sqrt spread
    std::uint32_t* pathScene;  // batch renderer: scene of each path
    // Subsurface walks deferred by shade, indexed by walk-queue slot: three
    // float4 per walk. In: (entry xyz, radius), (outward normal xyz, albedo),
    // (anisotropy, rng bits, -, -). Out: (exit origin xyz, attenuation),
    // (exit direction xyz, 1 = exited / 0 = no exit), (-, rng bits, -, -).
    float4* walk;
    std::uint32_t* walkQueue;  // path id of each walk slot
    std::uint32_t capacity;
};

// Scene of a path: one scene for the whole launch (held by value in the
// kernel arguments), or an entry of a device table of scenes selected per
// path (the batch renderer).
struct SingleScene {
    DeviceScene s;
    __device__ __forceinline__ const DeviceScene& Get(std::uint32_t) const { return s; }
};
struct SceneTable {
    const DeviceScene* table;
    const std::uint32_t* pathScene;
    __device__ __forceinline__ const DeviceScene& Get(std::uint32_t pid) const { return table[pathScene[pid]]; }
};

struct RcQuery;

// Radiance-cache hooks of the integrator. Training pixel samples (all three
// lanes) are those with (ps + trainOffset) % trainStride == 0; they run the
// plain estimator and record their first recordVertices vertices. Other
// paths stop at the first vertex from terminateBounce on whose path spread
// exceeds spreadThreshold x the primary footprint and queue a cache query
// (Mueller et al., "Real-time Neural Radiance Caching for Path Tracing",
// SIGGRAPH 2021, section 3.4: a(x1..xn) = (sum_i sqrt(|x_{i-1} - x_i|^2 /
// (p(w_i) |cos theta_i|)))^2 against a0 = |x0 - x1|^2 / (4 pi cos theta_1)).
struct CacheParams {
    std::uint32_t enabled;
    std::uint32_t terminateBounce;
    float spreadThreshold;         // <= 0: stop at terminateBounce unconditionally
    std::uint32_t trainStride;
    std::uint32_t trainOffset;
    std::uint32_t recordVertices;
    std::uint32_t queryCapacity;
    std::uint32_t recordSlots;     // training pixel-sample slots in the record buffer
    std::uint32_t sampleCapacity;  // training samples kept per frame (per scene in a batch)
    std::uint32_t reserved;        // keeps the struct free of implicit padding
    RcQuery* queries;              // per chunk
    std::uint32_t* queryCount;
    RcQuery* records;              // [slot][lane][vertex]; target holds the radiance after emission
    float* recordThroughput;
    RcQuery* samples;              // per frame
    std::uint32_t* sampleCount;
    // Batch renderer (null in a single render): record slots are handed out
    // per lane path at raygen (trainSlot[path], ~0 = not training; the
    // records are [slot][vertex]), training pixel samples are chosen per
    // scene, and samples go to the path's scene: samples[scene *
    // sampleCapacity + k], k counted in sceneSampleCount[scene].
    std::uint32_t* trainSlot;
    std::uint32_t* recordCount;    // slots handed out in the current wave
    const uint2* sceneTrain;       // per scene: (trainStride, trainOffset); stride 0 = no training
    std::uint32_t* sceneSampleCount;
};

// Pixel-sample chunk: samples [s0, s1) x pixels [p0, p1); path id =
// ((s - s0) * (p1 - p0) + (p - p0)) * 3 + lane.
struct ChunkDesc {
    std::uint32_t width, height;
    std::uint32_t sampleBase;        // sample number of s = 0
    std::uint32_t s0, s1, p0, p1;
    std::uint32_t maxBounces;
    float cam[12];                   // origin, lowerLeft - origin, horizontal, vertical
};

// Batch renderer: one view per scene, and tiles (scene, pixel range, all of
// the view's samples) packed into waves of at most the path-pool capacity.
// Lane paths of a tile: pathOffset + ((s * (p1 - p0) + (p - p0)) * 3 + lane).
struct BatchView {
    float cam[12];               // origin, lowerLeft - origin, horizontal, vertical
    std::uint32_t width, height;
    std::uint32_t sampleBase;    // sample number of the first sample of the frame
    std::uint32_t samples;       // samples per pixel in this frame
    double4* batch;              // FP64 per-pixel sums of this frame's samples
};
struct BatchTile {
    std::uint32_t scene;
    std::uint32_t p0, p1;
    std::uint32_t pathOffset;    // within its wave
    std::uint32_t pixelOffset;   // within its wave
    std::uint32_t pad[3];
};
struct BatchWaves {
    const BatchView* views;
    const BatchTile* tiles;          // wave by wave
    const std::uint32_t* waveTiles;  // wave w: tiles [waveTiles[w], waveTiles[w + 1])
    const std::uint32_t* wavePaths;  // lane paths of wave w
    const std::uint32_t* wavePixels; // tile pixels of wave w
    std::uint32_t* cursor;           // [0]: the wave being rendered
};

struct TraversalCounters {
    unsigned long long nodes;
    unsigned long long triangles;
    unsigned long long rays;
    unsigned long long hits;  // extend: paths given a surface hit (followers included)
};

// ---------------------------------------------------------------------------
// Small float3 algebra (explicit, so the contraction rules are visible).
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(float s) { return V3{s, s, s}; }
__device__ __forceinline__ V3 v3(const float4& a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 v3(const F4& a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a / length(a); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float sqf(float x) { return x * x; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ V3 reflect3(V3 i, V3 n) { return i - n * (2.0f * dot(n, i)); }
__device__ __forceinline__ V3 vmax(V3 a, float s) { return {fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }

// lowbias32 integer hash and its FP32 uniform (reference
// src/gpu/vulkan_path_tracer.cpp:100-113); bit-exact with any IEEE host.
__device__ __forceinline__ std::uint32_t hashState(std::uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352dU;
    v ^= v >> 15;
    v *= 0x846ca68bU;
    return v ^ (v >> 16);
}

__device__ __forceinline__ float randomFloat(std::uint32_t& state)
{
    state = hashState(state);
    return __uint2float_rn(state) * (1.0f / 4294967296.0f);
}

} // namespace hdcx
