// Device-side state shared by the single-scene tracer and the batch
// renderer: one scene's packed arrays and BVH (SceneStore), and the
// wavefront path state (PathPool).
#pragma once

#include "../../include/hdcx.h"
#include "bvh.hpp"
#include "dev_common.h"
#include "device_buffer.h"
#include "lbvh.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hdcx {

struct SceneStore {
    DeviceScene scene{};
    bool geometry = false;
    DeviceBuffer nodes, tris, positions, indices, triMaterial, texcoords, normals, emissive, shading, materials, lights,
        textures, texels, importance, importancePool;
    // Light tables built on the device: per-material emission power, FP64
    // scan workspace, one success flag per dome light.
    DeviceBuffer materialPower, tableScratch, domeOk;
    // Host copies for inspection: entries of the emissive table, each
    // light's importance layout (width 0: sampled uniformly).
    std::uint32_t emissiveEntries = 0;
    std::vector<ImportanceDesc> importanceLayout;
    // BVH: builder policy, the device builder's workspace, refit order and
    // the triangles (vertex count, index buffer) the BVH was built over.
    hdcx_bvh_options bvhOptions{HDCX_BVH_DEVICE_LBVH, 2048U, 1U, 0U};
    hdcx_bvh_info bvhInfo{};
    DeviceBvhBuilder lbvh;
    RefitPlan refitPlan;
    std::vector<std::uint32_t> builtIndices;
    std::uint32_t builtVertexCount = 0;
    // Two-level scenes: instance records, prototypes, flattened bases and
    // the object-space padding terms of each instance.
    DeviceBuffer instanceRecords, prototypeRecords, instanceBase, instancePad;
    // Vertex bounds (radiance-cache position normalisation).
    float sceneLo[3] = {0, 0, 0}, sceneInvExtent[3] = {1, 1, 1};
};

// The checks of hdcx_set_scene (throws Failure HDCX_ERR_INVALID_SCENE).
void ValidateScene(const hdcx_scene_view& v);

// Uploads a scene to the current device: validates it, uploads every
// packed array and builds the BVH, or refits it when the triangles are the
// same as last time and refit is on. A null or empty view leaves no
// geometry. Synchronises st.
void UploadScene(SceneStore* s, const hdcx_scene_view* view, hipStream_t st);

struct PathPool {
    DeviceBuffer rayO, rayD, hit, rng, wavelength, throughput, radiance, flags, mis, medium, shadowO, shadowD,
        shadowLight, queue0, queue1, shadowQueue, lambda0, spread, pathScene, walk, walkQueue, extendQueue0, extendQueue1,
        follow, shadowLeaders, shadowFollow, counters, totals;
    std::uint32_t capacity = 0;
    std::uint32_t have = 0;  // optional arrays allocated at `capacity` (kNeed* bits)

    // Optional per-path arrays, allocated only for renders that use them:
    // deferred subsurface walks (52 B per path), the radiance cache's path
    // spread (8 B), the batch renderer's scene index (4 B).
    static constexpr std::uint32_t kNeedWalks = 1U, kNeedSpread = 2U, kNeedSceneIndex = 4U;

    // Room for `paths` lane paths (grows only) and the optional arrays in
    // `needs`.
    void EnsurePaths(std::uint32_t paths, std::uint32_t needs);
    PathBuffers Paths();
};

} // namespace hdcx
