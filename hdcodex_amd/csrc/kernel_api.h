// Host-side launchers of the integrator kernels (defined in kernels.hip).
#pragma once

#include "dev_common.h"

namespace hdcx {

hipError_t LaunchRaygen(const PathBuffers& pb, const ChunkDesc& c, std::uint32_t paths, const CacheParams& cp,
                        unsigned maxBlocks, hipStream_t st);
// Training samples from the chunk's records; then the records are
// consumed and, with `outputs` (the cache's three outputs for every record
// query), training paths that reached the terminate vertex take the cached
// value there in the image, scaled by *calibration when given. With
// `scratch` (GatherScratchWords(records) words; one scene, records counted
// on the host) the samples follow record order; otherwise their order is
// that of the appends.
hipError_t LaunchGatherTraining(const PathBuffers& pb, const CacheParams& cp, std::uint32_t records, unsigned maxBlocks,
                                hipStream_t st, const float* outputs = nullptr, const float* calibration = nullptr,
                                std::uint32_t* scratch = nullptr);
std::size_t GatherScratchWords(std::uint32_t records);
// Extend sharing: kShareNone traces every queued path; kShareCamera (the
// launch after raygen) traces one camera ray per pixel sample for its
// three lanes; kShareLanes traces the paths the previous shade queued as
// tracing (its lane pairing) and writes their hits to the later lanes of
// their sample with the same ray.
constexpr int kShareNone = 0, kShareCamera = 1, kShareLanes = 2;
hipError_t LaunchExtend(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int sharing, std::uint32_t maxWork,
                        unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st);
// walk: kShadeWalkInline runs subsurface walks inside shade;
// kShadeWalkDefer queues them (then LaunchWalk, then shade again with
// kShadeWalkResume over the walk queue: same results, the walks' traversal
// runs in a kernel of its own). kShadeSimple: the variant without coat,
// sheen, transmission, subsurface and translucent lobes, for scenes whose
// materials have none (DeviceScene::simpleMaterials); same images.
constexpr int kShadeWalkInline = 0, kShadeWalkDefer = 1, kShadeWalkResume = 2, kShadeSimple = 3;
// Ends the paths of queue[in] whose ray missed (environment added) and
// queues the others for shade; run between extend and shade.
hipError_t LaunchMiss(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, std::uint32_t maxWork,
                      unsigned maxBlocks, hipStream_t st);
// The first bounce's miss (queue 0 = the launch's lane paths in order, three
// per pixel sample with one camera ray): one thread per pixel sample, the
// same results as LaunchMiss. maxWork: the launch's lane paths.
hipError_t LaunchMissCamera(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st);
// share: shade also pairs the lanes of a pixel sample with equal next rays
// (the queue the next extend reads with kShareLanes) and with equal shadow
// rays (LaunchShadow with shared = true traces each group once).
hipError_t LaunchShade(const DeviceScene& s, const PathBuffers& pb, std::uint32_t in, int bounce, int last,
                       const CacheParams& cp, int walk, int share, std::uint32_t maxWork, unsigned maxBlocks,
                       hipStream_t st);
hipError_t LaunchWalk(const DeviceScene& s, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                      hipStream_t st);
hipError_t LaunchShadow(const DeviceScene& s, const PathBuffers& pb, bool shared, std::uint32_t maxWork, unsigned maxBlocks,
                        TraversalCounters* ctr, hipStream_t st);
// Batch renderer: the same stages over a table of scenes (scene of each
// path in pb.pathScene; twoLevel when any scene has instances, opaque when
// every scene is opaque), the waves' raygen / film reading the device wave
// cursor, and the cursor step. maxWork sizes the grid (as a wave's paths or
// pixels can only be known on the device).
hipError_t LaunchExtendBatch(const DeviceScene* table, bool twoLevel, bool opaque, const PathBuffers& pb,
                             std::uint32_t in, int sharing, std::uint32_t maxWork, unsigned maxBlocks,
                             TraversalCounters* ctr, hipStream_t st);
hipError_t LaunchMissBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t in, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st);
// LaunchMissCamera for a batch wave (its lane paths are three per pixel
// sample, each sample's lanes consecutive and in one scene).
hipError_t LaunchMissCameraBatch(const DeviceScene* table, const PathBuffers& pb, std::uint32_t maxWork, unsigned maxBlocks,
                                 hipStream_t st);
hipError_t LaunchShadeBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t in, int bounce,
                            int last, const CacheParams& cp, int walk, int share, std::uint32_t maxWork, unsigned maxBlocks,
                            hipStream_t st);
hipError_t LaunchWalkBatch(const DeviceScene* table, bool twoLevel, const PathBuffers& pb, std::uint32_t maxWork,
                           unsigned maxBlocks, hipStream_t st);
// The batch always shares rays (its shade runs with share = 1).
hipError_t LaunchShadowBatch(const DeviceScene* table, bool twoLevel, bool opaque, const PathBuffers& pb,
                             std::uint32_t maxWork, unsigned maxBlocks, TraversalCounters* ctr, hipStream_t st);
hipError_t LaunchRaygenBatch(const PathBuffers& pb, const BatchWaves& bw, const CacheParams& cp, std::uint32_t maxWork,
                             unsigned maxBlocks, hipStream_t st);
hipError_t LaunchFilmBatch(const PathBuffers& pb, const BatchWaves& bw, std::uint32_t maxWork, unsigned maxBlocks,
                           hipStream_t st);
// Next wave; also zeroes *clear when given (the wave's record-slot count).
hipError_t LaunchAdvanceWave(const BatchWaves& bw, std::uint32_t* clear, hipStream_t st);
hipError_t LaunchFilm(const PathBuffers& pb, const ChunkDesc& c, double4* batch, unsigned maxBlocks, hipStream_t st);
// Counts the answered cache queries of a bounce and empties the queue.
hipError_t LaunchConsumeQueries(const CacheParams& cp, hipStream_t st);
// Marks n training records empty (before a single render's chunk).
hipError_t LaunchClearRecords(RcQuery* records, std::uint32_t n, unsigned maxBlocks, hipStream_t st);
hipError_t LaunchResolve(double4* mean, const double4* batch, float4* out, std::uint32_t pixels, std::uint32_t sampleIndex,
                         std::uint32_t sampleCount, unsigned maxBlocks, hipStream_t st);
// Nearest-neighbour upscale of a w0 x h0 RGBA32F film to w x h
// (src/hydra/render_pass.cpp:45-68).
hipError_t LaunchUpscaleNearest(const float4* src, std::uint32_t sw, std::uint32_t sh, float4* dst, std::uint32_t w,
                                std::uint32_t h, unsigned maxBlocks, hipStream_t st);
hipError_t LaunchZero(double4* p, std::uint32_t n, unsigned maxBlocks, hipStream_t st);
hipError_t LaunchQueryClosest(const DeviceScene& s, const float4* rays, std::uint32_t n, float4* hits, unsigned maxBlocks,
                              hipStream_t st);
hipError_t LaunchQueryAny(const DeviceScene& s, const float4* rays, std::uint32_t n, unsigned char* out, unsigned maxBlocks,
                          hipStream_t st);

} // namespace hdcx
