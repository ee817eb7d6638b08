// C ABI of the hdcx path-tracing core: tracer lifetime, scene upload, the
// progressive render loop, ray queries, procedural scenes and measurement.
// Errors never cross the ABI as exceptions: every entry point maps them to
// a status code and a thread-local message.
#include "../../include/hdcx.h"
#include "dev_traverse.h"
#include "device_buffer.h"
#include "scene_store.h"
#include "kernel_api.h"
#include "lbvh.h"
#include "lbvh_host.hpp"
#include "packer.hpp"
#include "procedural.hpp"
#include "rc_config.h"
#include "rcache.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace hdcx {
namespace {

constexpr std::uint32_t kChunkPixelSamples = 1U << 25;  // 100.7M lane paths (~29 GB of path state) per chunk

} // namespace
} // namespace hdcx

using namespace hdcx;

struct hdcx_scene {
    ProceduralScene procedural;
    PackedScene packed;
};

// The scene copy (SceneStore) and the wavefront path state (PathPool) are
// shared with the batch renderer.
struct hdcx_tracer : SceneStore, PathPool {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned maxBlocks = 1024;
    int shadingMode = HDCX_SHADING_FUSED;
    bool fallbackInit = false;
    bool fallbackZUp = false;
    DeviceBuffer traversalCounters;
    // Film.
    DeviceBuffer filmMean, filmBatch, filmOut, filmUpscaled;
    std::uint32_t filmWidth = 0, filmHeight = 0;
    // Ray queries.
    DeviceBuffer queryRays, queryHits;
    // Radiance cache.
    hdcx_radiance_cache_config rcConfig{};
    bool rcEnabled = false;
    bool rcInitialised = false;
    std::uint64_t rcSeed = 0;
    int rcStep = 0;
    hdcx_rc_stats rcStats{};
    DeviceBuffer rcMaster, rcM, rcV, rcGrad, rcW, rcWt, rcLoss, rcQueries, rcOut, rcEma, rcWe, rcWte, rcAct;
    DeviceBuffer rcChunkQueries, rcCounters, rcRecords, rcRecordThroughput, rcSamples, rcCalib, rcCalibOut, rcRecordOut,
        rcGather;
    std::uint32_t rcFrame = 0;
    bool rcCountersReady = false;
    // Instrumentation.
    std::uint32_t instrumentation = 0;
    std::uint32_t sampleOffset = 0;  // added to sample numbers in the random seeds only
    hdcx_kernel_stats stats{};
    std::vector<hipEvent_t> eventPool;
    struct Timed { int kernel; hipEvent_t a, b; };
    std::vector<Timed> pending;
    std::size_t eventsUsed = 0;

    ~hdcx_tracer()
    {
        for (hipEvent_t e : eventPool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void Activate() const { Check(hipSetDevice(device), "hipSetDevice"); }

    hipEvent_t NextEvent()
    {
        if (eventsUsed == eventPool.size()) {
            hipEvent_t e;
            Check(hipEventCreate(&e), "hipEventCreate");
            eventPool.push_back(e);
        }
        return eventPool[eventsUsed++];
    }

    template <class Launch>
    void Timed(int kernel, Launch&& launch)
    {
        hipEvent_t a = nullptr, b = nullptr;
        if (instrumentation & 1U) {
            a = NextEvent();
            b = NextEvent();
            Check(hipEventRecord(a, stream), "hipEventRecord");
        }
        Check(launch(), "kernel launch");
        if (instrumentation & 1U) {
            Check(hipEventRecord(b, stream), "hipEventRecord");
            pending.push_back({kernel, a, b});
        }
        ++stats.launches[kernel];
    }

    void CollectTimings()
    {
        for (const auto& t : pending) {
            float ms = 0.0F;
            Check(hipEventElapsedTime(&ms, t.a, t.b), "hipEventElapsedTime");
            stats.milliseconds[t.kernel] += ms;
        }
        pending.clear();
        eventsUsed = 0;
    }

    RcState Rc()
    {
        RcState r{};
        r.master = rcMaster.as<float>();
        r.m = rcM.as<float>();
        r.v = rcV.as<float>();
        r.grad = rcGrad.as<float>();
        r.w = rcW.as<__bf16>();
        r.wt = rcWt.as<__bf16>();
        r.ema = rcEma.as<float>();
        r.we = rcWe.as<__bf16>();
        r.wte = rcWte.as<__bf16>();
        r.emaDecay = rcConfig.ema_decay;
        r.loss = rcLoss.as<float>();
        return r;
    }

    // Device address of the brightness calibration factor, or null when
    // calibration is off.
    const float* CalibrationScale()
    {
        if (!(rcConfig.calibration_decay > 0.0F) || !rcCalib.bytes()) return nullptr;
        return reinterpret_cast<const float*>(rcCalib.as<char>() + offsetof(RcCalibration, scale));
    }

    // Weights of the training passes, and the (averaged) inference weights.
    RcNetwork RcNet() { return RcNetwork{rcW.as<__bf16>(), rcWt.as<__bf16>()}; }
    RcNetwork RcInferNet() { return RcNetwork{rcWe.as<__bf16>(), rcWte.as<__bf16>()}; }

    // Allocates the cache state and (re)initialises the weights for `seed`:
    // He-uniform bounds sqrt(6 / fan_in), the output layer scaled by 0.5;
    // Adam moments and gradient partials zeroed.
    void EnsureRc(std::uint64_t seed)
    {
        if (rcInitialised && seed == rcSeed) return;
        Activate();
        rcMaster.Allocate(kRcParams * sizeof(float));
        rcM.Allocate(kRcParams * sizeof(float));
        rcV.Allocate(kRcParams * sizeof(float));
        rcGrad.Allocate(static_cast<std::size_t>(kRcGradSplits) * kRcParams * sizeof(float));
        rcW.Allocate(kRcParams * sizeof(std::uint16_t));
        rcWt.Allocate(4 * 4096 * sizeof(std::uint16_t));
        rcEma.Allocate(kRcParams * sizeof(float));
        rcWe.Allocate(kRcParams * sizeof(std::uint16_t));
        rcWte.Allocate(4 * 4096 * sizeof(std::uint16_t));
        rcLoss.Allocate(2 * sizeof(float));
        rcCalib.Allocate(sizeof(RcCalibration));
        const RcCalibration calib{{0.0, 0.0}, {0.0, 0.0}, 1.0F, {0.0F, 0.0F, 0.0F}};
        Check(hipMemcpyAsync(rcCalib.as<void>(), &calib, sizeof(calib), hipMemcpyHostToDevice, stream), "rc init");
        std::vector<float> w(kRcParams);
        InitRcWeights(seed, w.data());
        Check(hipMemcpyAsync(rcMaster.as<void>(), w.data(), kRcParams * sizeof(float), hipMemcpyHostToDevice, stream), "rc init");
        Check(hipMemsetAsync(rcM.as<void>(), 0, rcM.bytes(), stream), "rc init");
        Check(hipMemsetAsync(rcV.as<void>(), 0, rcV.bytes(), stream), "rc init");
        Check(hipMemsetAsync(rcGrad.as<void>(), 0, rcGrad.bytes(), stream), "rc init");
        Check(LaunchRcPublish(Rc(), stream), "rc publish");
        Check(hipStreamSynchronize(stream), "rc init");
        rcSeed = seed;
        rcStep = 0;
        rcInitialised = true;
    }

    // One Adam step per mini-batch over `count` samples already on the device.
    float RcTrain(const RcQuery* samples, const std::uint32_t* deviceCount, std::uint32_t count, std::uint32_t steps)
    {
        const hdcx_radiance_cache_config& c = rcConfig;
        steps = std::max(steps, 1U);
        rcAct.Allocate(std::max(rcAct.bytes(), RcGradWorkspaceBytes(count / steps + 1U)));
        float lossHost[2] = {0.0F, 0.0F};
        for (std::uint32_t i = 0; i < steps; ++i) {
            const std::uint32_t b0 = static_cast<std::uint32_t>(static_cast<std::uint64_t>(count) * i / steps);
            const std::uint32_t b1 = static_cast<std::uint32_t>(static_cast<std::uint64_t>(count) * (i + 1) / steps);
            if (b1 <= b0) continue;
            (void)deviceCount;
            Check(hipMemsetAsync(rcLoss.as<void>(), 0, 2 * sizeof(float), stream), "rc loss");
            Check(LaunchRcGradient(RcNet(), samples + b0, b1 - b0, rcAct.as<__bf16>(), rcGrad.as<float>(),
                                   rcLoss.as<float>(), stream),
                  "rc gradient");
            Check(LaunchRcAdamBlocks(Rc(), kRcGradSplits, ++rcStep, c.learning_rate, c.beta1, c.beta2,
                                     c.epsilon, stream),
                  "rc adam");
            ++rcStats.adam_steps;
            rcStats.train_samples += b1 - b0;
        }
        Check(hipMemcpyAsync(lossHost, rcLoss.as<void>(), sizeof(lossHost), hipMemcpyDeviceToHost, stream), "rc loss");
        Check(hipStreamSynchronize(stream), "rc train");
        rcStats.last_loss = lossHost[1] > 0.0F ? lossHost[0] / lossHost[1] : 0.0F;
        return rcStats.last_loss;
    }

    void EnsureFilm(std::uint32_t w, std::uint32_t h)
    {
        if (w == filmWidth && h == filmHeight && filmMean.bytes()) return;
        const std::size_t n = static_cast<std::size_t>(w) * h;
        filmMean.Allocate(n * sizeof(double4));
        filmBatch.Allocate(n * sizeof(double4));
        filmOut.Allocate(n * sizeof(float4));
        filmWidth = w;
        filmHeight = h;
        // A resized film starts from zero (reference EnsureOutput :3130-3160).
        Check(LaunchZero(filmMean.as<double4>(), static_cast<std::uint32_t>(n), maxBlocks, stream), "zero film");
    }
};

namespace hdcx {
namespace {

void SetScene(hdcx_tracer* t, const hdcx_scene_view* view)
{
    t->Activate();
    UploadScene(t, view, t->stream);
    t->stats.node_bytes = sizeof(Bvh4Node);
    t->stats.triangle_bytes = 48;
    t->stats.bvh_nodes = t->bvhInfo.node_count;
    t->stats.bvh_triangles = t->bvhInfo.slot_count;
}

void Render(hdcx_tracer* t, const hdcx_camera& cam, std::uint32_t w, std::uint32_t h, std::uint32_t sampleIndex,
            std::uint32_t maxBounces, std::uint32_t sampleCount, float* out)
{
    t->Activate();
    if (!t->fallbackInit) {
        // Fallback sky axis from the first camera (:3237-3241).
        t->fallbackZUp = std::fabs(cam.vertical[2]) > std::fabs(cam.vertical[1]);
        t->fallbackInit = true;
    }
    t->scene.zUp = t->fallbackZUp ? 1U : 0U;
    const std::uint32_t samples = std::max(sampleCount, 1U);
    const std::uint32_t bounces = std::clamp(maxBounces, 1U, kMaxBounces);
    const std::uint32_t pixels = w * h;
    t->EnsureFilm(w, h);
    hipStream_t st = t->stream;
    Check(LaunchZero(t->filmBatch.as<double4>(), pixels, t->maxBlocks, st), "zero batch");
    if (!t->traversalCounters.bytes()) {
        t->traversalCounters.Allocate(sizeof(TraversalCounters) * 2);
        Check(hipMemsetAsync(t->traversalCounters.as<void>(), 0, sizeof(TraversalCounters) * 2, st), "memset counters");
    }
    // Chunking: whole frames of k samples when a frame is small, else pixel
    // ranges of one sample; no pixel appears twice within a chunk.
    std::uint32_t pixelsPerChunk = std::min(pixels, kChunkPixelSamples);
    std::uint32_t samplesPerChunk = pixels <= kChunkPixelSamples ? std::max(1U, kChunkPixelSamples / pixels) : 1U;
    samplesPerChunk = std::min(samplesPerChunk, samples);
    {
        // The same number of samples in every chunk (to within one): a short
        // last chunk would run its launches on a fraction of the GPU.
        const std::uint32_t chunks = (samples + samplesPerChunk - 1U) / samplesPerChunk;
        samplesPerChunk = (samples + chunks - 1U) / chunks;
    }
    const bool deferWalks = t->scene.hasSubsurface && !(t->instrumentation & 4U);
    t->EnsurePaths(pixelsPerChunk * samplesPerChunk * 3U,
                   (deferWalks ? PathPool::kNeedWalks : 0U) | (t->rcEnabled ? PathPool::kNeedSpread : 0U));
    PathBuffers pb = t->Paths();
    // Radiance cache: training pixel samples are spread evenly over the
    // frame so that about train_samples vertices are recorded (2.5 per lane
    // path on average, 3 lanes, up to 4 vertices each).
    CacheParams cp{};
    const bool cache = t->rcEnabled;
    constexpr std::uint32_t kRecordVertices = 4;
    if (cache) {
        const std::uint64_t frameSamples = static_cast<std::uint64_t>(pixels) * samples;
        const std::uint64_t wanted = std::max<std::uint64_t>(1, t->rcConfig.train_samples * 2ULL / 15ULL);
        cp.enabled = 1;
        cp.terminateBounce = t->rcConfig.terminate_bounce;
        cp.spreadThreshold = t->rcConfig.spread_threshold;
        cp.trainStride = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, frameSamples / wanted));
        cp.trainOffset = static_cast<std::uint32_t>((t->rcFrame * 2654435761ULL) % cp.trainStride);
        cp.recordVertices = kRecordVertices;
        cp.queryCapacity = t->capacity;
        cp.recordSlots = pixelsPerChunk * samplesPerChunk / cp.trainStride + 2;
        cp.sampleCapacity = t->rcConfig.train_samples;
        t->rcChunkQueries.Allocate(std::max<std::size_t>(t->rcChunkQueries.bytes(), static_cast<std::size_t>(t->capacity) * sizeof(RcQuery)));
        t->rcCounters.Allocate(64);
        const std::size_t records = static_cast<std::size_t>(cp.recordSlots) * 3 * kRecordVertices;
        t->rcRecords.Allocate(std::max<std::size_t>(t->rcRecords.bytes(), records * sizeof(RcQuery)));
        t->rcRecordThroughput.Allocate(std::max<std::size_t>(t->rcRecordThroughput.bytes(), records * sizeof(float)));
        t->rcSamples.Allocate(std::max<std::size_t>(t->rcSamples.bytes(), static_cast<std::size_t>(cp.sampleCapacity) * sizeof(RcQuery)));
        cp.queries = t->rcChunkQueries.as<RcQuery>();
        cp.queryCount = t->rcCounters.as<std::uint32_t>();
        cp.records = t->rcRecords.as<RcQuery>();
        cp.recordThroughput = t->rcRecordThroughput.as<float>();
        cp.samples = t->rcSamples.as<RcQuery>();
        cp.sampleCount = t->rcCounters.as<std::uint32_t>() + 1;
        if (!t->rcCountersReady) {
            Check(hipMemsetAsync(t->rcCounters.as<void>(), 0, 64, st), "rc counters");
            t->rcCountersReady = true;
        }
        Check(hipMemsetAsync(t->rcCounters.as<void>(), 0, 8, st), "rc counters");
    }
    TraversalCounters* ctr = (t->instrumentation & 2U) ? t->traversalCounters.as<TraversalCounters>() : nullptr;
    const int walk = t->scene.simpleMaterials && !(t->instrumentation & 16U) ? kShadeSimple
                     : deferWalks                                             ? kShadeWalkDefer
                                                                              : kShadeWalkInline;
    const bool share = !(t->instrumentation & 8U);  // lanes of a pixel sample trace an equal ray once
    TraversalCounters* shadowCtr = ctr ? ctr + 1 : nullptr;
    ChunkDesc c{};
    c.width = w;
    c.height = h;
    c.sampleBase = sampleIndex + t->sampleOffset;
    c.maxBounces = bounces;
    std::memcpy(c.cam + 0, cam.origin, 12);
    for (int k = 0; k < 3; ++k)
        c.cam[3 + k] = static_cast<float>(static_cast<double>(cam.lower_left[k]) - static_cast<double>(cam.origin[k]));
    std::memcpy(c.cam + 6, cam.horizontal, 12);
    std::memcpy(c.cam + 9, cam.vertical, 12);
    for (std::uint32_t s0 = 0; s0 < samples; s0 += samplesPerChunk) {
        const std::uint32_t s1 = std::min(samples, s0 + samplesPerChunk);
        for (std::uint32_t p0 = 0; p0 < pixels; p0 += pixelsPerChunk) {
            const std::uint32_t p1 = std::min(pixels, p0 + pixelsPerChunk);
            c.s0 = s0;
            c.s1 = s1;
            c.p0 = p0;
            c.p1 = p1;
            const std::uint32_t paths = (s1 - s0) * (p1 - p0) * 3U;
            if (cache)
                Check(LaunchClearRecords(cp.records, cp.recordSlots * 3U * kRecordVertices, t->maxBlocks, st),
                      "clear training records");
            t->Timed(HDCX_KERNEL_RAYGEN, [&] { return LaunchRaygen(pb, c, paths, cp, t->maxBlocks, st); });
            t->stats.rays[HDCX_KERNEL_RAYGEN] += paths;
            for (std::uint32_t b = 0; b < bounces; ++b) {
                const std::uint32_t in = b & 1U;
                const int sharing = !share ? kShareNone : b == 0 ? kShareCamera : kShareLanes;
                t->Timed(HDCX_KERNEL_EXTEND,
                         [&] { return LaunchExtend(t->scene, pb, in, sharing, paths, t->maxBlocks, ctr, st); });
                const int last = b + 1 == bounces ? 1 : 0;
                t->Timed(HDCX_KERNEL_MISS, [&] {
                    return b == 0 ? LaunchMissCamera(t->scene, pb, paths, t->maxBlocks, st)
                                  : LaunchMiss(t->scene, pb, in, paths, t->maxBlocks, st);
                });
                // With sharing on, shade also pairs the lanes of each pixel
                // sample with equal next rays and equal shadow rays.
                t->Timed(HDCX_KERNEL_SHADE, [&] {
                    return LaunchShade(t->scene, pb, in, static_cast<int>(b), last, cp, walk, share ? 1 : 0, paths,
                                       t->maxBlocks, st);
                });
                if (walk == kShadeWalkDefer && !last) {
                    t->Timed(HDCX_KERNEL_WALK, [&] { return LaunchWalk(t->scene, pb, paths, t->maxBlocks, st); });
                    t->Timed(HDCX_KERNEL_SHADE, [&] {
                        return LaunchShade(t->scene, pb, in, static_cast<int>(b), last, cp, kShadeWalkResume, share ? 1 : 0,
                                           paths, t->maxBlocks, st);
                    });
                }
                // Queries come from the terminate vertex, or from any later
                // one when the spread rule is on.
                if (cache && (b == cp.terminateBounce || (b > cp.terminateBounce && cp.spreadThreshold > 0.0F))) {
                    // Queries of this bounce: answered, then the queue is emptied.
                    t->Timed(HDCX_KERNEL_CACHE, [&] {
                        return LaunchRcInfer(t->RcInferNet(), cp.queries, cp.queryCount, paths, nullptr, pb.radiance, pb.throughput,
                                             t->maxBlocks, st, t->CalibrationScale());
                    });
                    Check(LaunchConsumeQueries(cp, st), "rc queries");
                }
                t->Timed(HDCX_KERNEL_SHADOW,
                         [&] { return LaunchShadow(t->scene, pb, share, paths, t->maxBlocks, shadowCtr, st); });
            }
            if (cache) {
                const std::uint32_t records = cp.recordSlots * 3U * kRecordVertices;
                // Training paths show the cached value in the image too
                // (their plain rest of the path only trains the cache),
                // unless asked to keep their plain radiance or the stop
                // vertex depends on the spread rule.
                const bool cachedImage = !(t->rcConfig.flags & HDCX_RC_PLAIN_TRAINING_PATHS) && !(cp.spreadThreshold > 0.0F);
                const float* outputs = nullptr;
                if (cachedImage) {
                    t->rcRecordOut.Reserve(static_cast<std::size_t>(records) * 3 * sizeof(float));
                    Check(LaunchRcInfer(t->RcInferNet(), cp.records, nullptr, records, t->rcRecordOut.as<float>(), nullptr,
                                        nullptr, t->maxBlocks, st),
                          "rc training-path answers");
                    outputs = t->rcRecordOut.as<float>();
                }
                t->rcGather.Reserve(GatherScratchWords(records) * sizeof(std::uint32_t));
                Check(LaunchGatherTraining(pb, cp, records, t->maxBlocks, st, outputs, t->CalibrationScale(),
                                           t->rcGather.as<std::uint32_t>()),
                      "gather training samples");
            }
            t->Timed(HDCX_KERNEL_FILM, [&] { return LaunchFilm(pb, c, t->filmBatch.as<double4>(), t->maxBlocks, st); });
            t->stats.rays[HDCX_KERNEL_FILM] += (p1 - p0);
        }
    }
    Check(LaunchResolve(t->filmMean.as<double4>(), t->filmBatch.as<double4>(), t->filmOut.as<float4>(), pixels,
                        sampleIndex, samples, t->maxBlocks, st),
          "resolve");
    if (out) Check(hipMemcpyAsync(out, t->filmOut.as<void>(), static_cast<std::size_t>(pixels) * 16, hipMemcpyDeviceToHost, st),
                   "film readback");
    Check(hipStreamSynchronize(st), "render");
    if (cache) {
        // Online training on this frame's samples; the next frame uses the
        // updated weights.
        std::uint32_t counts[2] = {0, 0};
        Check(hipMemcpy(counts, t->rcCounters.as<void>(), sizeof(counts), hipMemcpyDeviceToHost), "rc counters");
        const std::uint32_t n = std::min(counts[1], cp.sampleCapacity);
        hipEvent_t a = nullptr, b = nullptr;
        if (t->instrumentation & 1U) {
            a = t->NextEvent();
            b = t->NextEvent();
            Check(hipEventRecord(a, st), "hipEventRecord");
        }
        if (n > 0) {
            t->RcTrain(cp.samples, nullptr, n, t->rcConfig.train_steps);
            if (t->CalibrationScale()) {
                // The updated inference network on this frame's samples
                // against their plain-estimator targets.
                t->rcCalibOut.Allocate(std::max<std::size_t>(t->rcCalibOut.bytes(), static_cast<std::size_t>(n) * 3 * sizeof(float)));
                Check(LaunchRcInfer(t->RcInferNet(), cp.samples, nullptr, n, t->rcCalibOut.as<float>(), nullptr, nullptr,
                                    t->maxBlocks, st),
                      "rc calibration outputs");
                Check(LaunchRcCalibrate(cp.samples, n, t->rcCalibOut.as<float>(), t->rcCalib.as<RcCalibration>(),
                                        t->rcConfig.calibration_decay, st),
                      "rc calibration");
                Check(hipStreamSynchronize(st), "rc calibration");
            }
        }
        if (t->instrumentation & 1U) {
            Check(hipEventRecord(b, st), "hipEventRecord");
            Check(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.0F;
            Check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            t->rcStats.train_ms += ms;
        }
        ++t->rcStats.frames;
        ++t->rcFrame;
    }
    t->CollectTimings();
}

void Query(hdcx_tracer* t, const hdcx_ray* rays, std::uint32_t count, void* out, bool closest)
{
    t->Activate();
    if (count == 0) return;
    hipStream_t st = t->stream;
    t->queryRays.Upload(rays, static_cast<std::size_t>(count) * sizeof(hdcx_ray), st);
    t->queryHits.Allocate(static_cast<std::size_t>(count) * (closest ? sizeof(hdcx_hit) : 1));
    if (closest)
        Check(LaunchQueryClosest(t->scene, t->queryRays.as<float4>(), count, t->queryHits.as<float4>(), t->maxBlocks, st), "query");
    else
        Check(LaunchQueryAny(t->scene, t->queryRays.as<float4>(), count, t->queryHits.as<unsigned char>(), t->maxBlocks, st), "query");
    Check(hipMemcpyAsync(out, t->queryHits.as<void>(), static_cast<std::size_t>(count) * (closest ? sizeof(hdcx_hit) : 1),
                         hipMemcpyDeviceToHost, st),
          "query readback");
    Check(hipStreamSynchronize(st), "query");
}

} // namespace
} // namespace hdcx

namespace {

void CopyBvhOut(const Bvh4& bvh, hdcx_bvh_info* out_info, void* out_nodes, float* out_slots)
{
    out_info->node_count = static_cast<std::uint32_t>(bvh.nodes.size());
    out_info->slot_count = static_cast<std::uint32_t>(bvh.triangles.size() / 12);
    out_info->max_stack = bvh.maxStack;
    out_info->depth = bvh.depth;
    if (out_nodes) std::memcpy(out_nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(Bvh4Node));
    if (out_slots) std::memcpy(out_slots, bvh.triangles.data(), bvh.triangles.size() * sizeof(float));
}

Bvh4 BuildOnDevice(const hdcx_scene_view& v, std::uint32_t topClusters, std::uint32_t& clusters)
{
    struct Stream {
        hipStream_t s = nullptr;
        ~Stream() { if (s) (void)hipStreamDestroy(s); }
    } stream;
    Check(hipStreamCreate(&stream.s), "hipStreamCreate");
    DeviceBuffer pos, idx, nodes, slots;
    pos.Upload(v.positions, static_cast<std::size_t>(v.vertex_count) * 16, stream.s);
    idx.Upload(v.indices, static_cast<std::size_t>(v.triangle_count) * 12, stream.s);
    float lo[3], hi[3];
    VertexBounds(v.positions, v.vertex_count, lo, hi);
    DeviceBvhBuilder builder;
    RefitPlan plan;
    DeviceBvhStats st;
    builder.Build(pos.as<float4>(), idx.as<std::uint32_t>(), v.triangle_count, lo, hi, topClusters, stream.s, nodes,
                  slots, plan, st);
    Bvh4 bvh;
    bvh.nodes.resize(st.nodeCount);
    bvh.triangles.resize(static_cast<std::size_t>(v.triangle_count) * 12);
    Check(hipMemcpy(bvh.nodes.data(), nodes.as<void>(), bvh.nodes.size() * sizeof(Bvh4Node), hipMemcpyDeviceToHost),
          "BVH download");
    Check(hipMemcpy(bvh.triangles.data(), slots.as<void>(), bvh.triangles.size() * sizeof(float), hipMemcpyDeviceToHost),
          "BVH download");
    bvh.maxStack = st.maxStack;
    bvh.depth = st.depth;
    clusters = st.topClusters;
    return bvh;
}

} // namespace

extern "C" {

const char* hdcx_last_error(void) { return hdcx::LastError().c_str(); }

int hdcx_create(int device_ordinal, hdcx_tracer** out_tracer)
{
    return Guard([&] {
        if (!out_tracer) throw std::invalid_argument("out_tracer is null");
        *out_tracer = nullptr;
        int devices = 0;
        Check(hipGetDeviceCount(&devices), "hipGetDeviceCount");
        if (device_ordinal < 0 || device_ordinal >= devices) throw Failure(HDCX_ERR_DEVICE, "no such HIP device");
        auto t = std::make_unique<hdcx_tracer>();
        t->device = device_ordinal;
        t->Activate();
        Check(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "hipStreamCreate");
        int cus = 0;
        Check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_ordinal), "device attribute");
        t->maxBlocks = static_cast<unsigned>(std::max(cus, 1)) * 16U;
        *out_tracer = t.release();
    });
}

int hdcx_destroy(hdcx_tracer* tracer)
{
    return Guard([&] {
        if (!tracer) return;
        (void)hipSetDevice(tracer->device);
        (void)hipStreamSynchronize(tracer->stream);
        delete tracer;
    });
}

int hdcx_set_scene(hdcx_tracer* tracer, const hdcx_scene_view* scene)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        SetScene(tracer, scene);
    });
}

int hdcx_set_sample_offset(hdcx_tracer* tracer, uint32_t offset)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        tracer->sampleOffset = offset;
    });
}

int hdcx_set_shading_mode(hdcx_tracer* tracer, int mode)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        if (mode == HDCX_SHADING_RASTER_PREVIEW)
            throw std::invalid_argument("raster preview is not a path-tracer shading mode");
        if (mode != HDCX_SHADING_FUSED && mode != HDCX_SHADING_MODULAR) throw std::invalid_argument("unknown shading mode");
        tracer->shadingMode = mode;
    });
}

int hdcx_get_shading_mode(const hdcx_tracer* tracer, int* out_mode)
{
    return Guard([&] {
        if (!tracer || !out_mode) throw std::invalid_argument("null argument");
        *out_mode = tracer->shadingMode;
    });
}

int hdcx_render(hdcx_tracer* tracer, const hdcx_camera* camera, uint32_t width, uint32_t height, uint32_t sample_index,
                uint32_t max_bounces, uint32_t sample_count, float* out_rgba)
{
    return Guard([&] {
        if (!tracer || !camera) throw std::invalid_argument("null argument");
        if (!tracer->geometry || width == 0 || height == 0) throw Failure(HDCX_ERR_NO_GEOMETRY, "no geometry or empty image");
        if (static_cast<std::uint64_t>(width) * height > 0x7fffffffULL) throw std::invalid_argument("image too large");
        Render(tracer, *camera, width, height, sample_index, max_bounces, sample_count, out_rgba);
    });
}

int hdcx_read_film(hdcx_tracer* tracer, float* out_rgba)
{
    return Guard([&] {
        if (!tracer || !out_rgba) throw std::invalid_argument("null argument");
        if (!tracer->filmOut.bytes()) throw Failure(HDCX_ERR_NO_GEOMETRY, "nothing rendered yet");
        tracer->Activate();
        Check(hipMemcpy(out_rgba, tracer->filmOut.as<void>(), tracer->filmOut.bytes(), hipMemcpyDeviceToHost), "readback");
    });
}

int hdcx_read_film_upscaled(hdcx_tracer* tracer, uint32_t width, uint32_t height, float* out_rgba, int out_on_device)
{
    return Guard([&] {
        if (!tracer || !out_rgba) throw std::invalid_argument("null argument");
        if (width == 0 || height == 0) throw std::invalid_argument("empty upscale target");
        if (static_cast<std::uint64_t>(width) * height >= (1ULL << 32) ||
            static_cast<std::uint64_t>(width) * tracer->filmWidth >= (1ULL << 32) ||
            static_cast<std::uint64_t>(height) * tracer->filmHeight >= (1ULL << 32))
            throw std::invalid_argument("upscale size beyond 32-bit index arithmetic");
        if (!tracer->filmOut.bytes()) throw Failure(HDCX_ERR_NO_GEOMETRY, "nothing rendered yet");
        tracer->Activate();
        const std::size_t bytes = static_cast<std::size_t>(width) * height * sizeof(float4);
        float4* dst = reinterpret_cast<float4*>(out_rgba);
        if (!out_on_device) {
            tracer->filmUpscaled.Reserve(bytes);
            dst = tracer->filmUpscaled.as<float4>();
        }
        Check(LaunchUpscaleNearest(tracer->filmOut.as<float4>(), tracer->filmWidth, tracer->filmHeight, dst, width, height,
                                   tracer->maxBlocks, tracer->stream),
              "upscale");
        if (!out_on_device)
            Check(hipMemcpyAsync(out_rgba, dst, bytes, hipMemcpyDeviceToHost, tracer->stream), "upscaled readback");
        Check(hipStreamSynchronize(tracer->stream), "upscale");
    });
}

int hdcx_has_geometry(const hdcx_tracer* tracer) { return tracer && tracer->geometry ? 1 : 0; }

int hdcx_trace_closest(hdcx_tracer* tracer, const hdcx_ray* rays, uint32_t count, hdcx_hit* out_hits)
{
    return Guard([&] {
        if (!tracer || (count && (!rays || !out_hits))) throw std::invalid_argument("null argument");
        if (!tracer->geometry) throw Failure(HDCX_ERR_NO_GEOMETRY, "no geometry");
        Query(tracer, rays, count, out_hits, true);
    });
}

int hdcx_trace_occluded(hdcx_tracer* tracer, const hdcx_ray* rays, uint32_t count, uint8_t* out_occluded)
{
    return Guard([&] {
        if (!tracer || (count && (!rays || !out_occluded))) throw std::invalid_argument("null argument");
        if (!tracer->geometry) throw Failure(HDCX_ERR_NO_GEOMETRY, "no geometry");
        Query(tracer, rays, count, out_occluded, false);
    });
}

int hdcx_scene_create_procedural(const char* name, uint64_t seed, uint32_t param, uint32_t param_dome_width,
                                 hdcx_scene** out_scene)
{
    return Guard([&] {
        if (!name || !out_scene) throw std::invalid_argument("null argument");
        auto s = std::make_unique<hdcx_scene>();
        s->procedural = MakeProceduralScene(name, seed, param, param_dome_width);
        s->packed = PackScene(s->procedural.snapshot);
        *out_scene = s.release();
    });
}

int hdcx_scene_view_get(const hdcx_scene* scene, hdcx_scene_view* out_view)
{
    return Guard([&] {
        if (!scene || !out_view) throw std::invalid_argument("null argument");
        *out_view = const_cast<hdcx_scene*>(scene)->packed.View();
    });
}

int hdcx_scene_flatten(const hdcx_scene* scene, hdcx_scene** out_scene)
{
    return Guard([&] {
        if (!scene || !out_scene) throw std::invalid_argument("null argument");
        auto s = std::make_unique<hdcx_scene>();
        s->procedural.eye = scene->procedural.eye;
        s->procedural.target = scene->procedural.target;
        s->procedural.up = scene->procedural.up;
        s->procedural.verticalFovDegrees = scene->procedural.verticalFovDegrees;
        s->procedural.explicitCamera = scene->procedural.explicitCamera;
        s->procedural.camera = scene->procedural.camera;
        s->packed = FlattenInstances(const_cast<hdcx_scene*>(scene)->packed.View());
        *out_scene = s.release();
    });
}

int hdcx_scene_camera_get(const hdcx_scene* scene, uint32_t width, uint32_t height, hdcx_camera* out_camera)
{
    return Guard([&] {
        if (!scene || !out_camera) throw std::invalid_argument("null argument");
        const PathTracerCamera c = scene->procedural.Camera(width, height);
        for (int k = 0; k < 3; ++k) {
            out_camera->origin[k] = c.origin[k];
            out_camera->lower_left[k] = c.lowerLeft[k];
            out_camera->horizontal[k] = c.horizontal[k];
            out_camera->vertical[k] = c.vertical[k];
        }
    });
}

int hdcx_scene_destroy(hdcx_scene* scene)
{
    delete scene;
    return HDCX_OK;
}

int hdcx_validate_scene(const hdcx_scene_view* scene)
{
    return Guard([&] {
        if (!scene) throw std::invalid_argument("scene is null");
        if (scene->vertex_count == 0 || scene->triangle_count == 0) return;
        ValidateScene(*scene);
    });
}

int hdcx_bvh_build_host(const hdcx_scene_view* scene, hdcx_bvh_info* out_info, void* out_nodes, float* out_slots)
{
    return hdcx_bvh_build(scene, HDCX_BVH_HOST_SAH, 0U, out_info, out_nodes, out_slots);
}

int hdcx_bvh_build(const hdcx_scene_view* scene, int builder, uint32_t top_clusters, hdcx_bvh_info* out_info,
                   void* out_nodes, float* out_slots)
{
    return Guard([&] {
        if (!scene || !out_info) throw std::invalid_argument("null argument");
        if (scene->triangle_count && (!scene->positions || !scene->indices)) throw std::invalid_argument("no geometry");
        if (builder != HDCX_BVH_HOST_SAH && builder != HDCX_BVH_DEVICE_LBVH && builder != HDCX_BVH_SERIAL_LBVH)
            throw std::invalid_argument("unknown BVH builder");
        if (builder != HDCX_BVH_HOST_SAH && scene->triangle_count < 2)
            throw std::invalid_argument("the LBVH builders need at least two triangles");
        for (std::size_t i = 0; i < static_cast<std::size_t>(scene->triangle_count) * 3; ++i)
            if (scene->indices[i] >= scene->vertex_count) throw std::invalid_argument("index out of range");
        const auto t0 = std::chrono::steady_clock::now();
        std::uint32_t clusters = 0;
        Bvh4 bvh;
        if (builder == HDCX_BVH_HOST_SAH) {
            bvh = BuildBvh4Host(scene->positions, scene->indices, scene->triangle_count);
        } else if (builder == HDCX_BVH_SERIAL_LBVH) {
            LbvhBuildInfo info;
            bvh = BuildBvh4LbvhSerial(scene->positions, scene->indices, scene->triangle_count, top_clusters, &info);
            clusters = info.topClusters;
        } else {
            bvh = BuildOnDevice(*scene, top_clusters, clusters);
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ValidateBvh4(bvh, scene->triangle_count);
        *out_info = hdcx_bvh_info{};
        CopyBvhOut(bvh, out_info, out_nodes, out_slots);
        out_info->builder = builder;
        out_info->top_clusters = clusters;
        out_info->build_ms = ms;
    });
}

int hdcx_set_bvh_options(hdcx_tracer* tracer, const hdcx_bvh_options* options)
{
    return Guard([&] {
        if (!tracer || !options) throw std::invalid_argument("null argument");
        if (options->builder != HDCX_BVH_HOST_SAH && options->builder != HDCX_BVH_DEVICE_LBVH &&
            options->builder != HDCX_BVH_SERIAL_LBVH)
            throw std::invalid_argument("unknown BVH builder");
        tracer->bvhOptions = *options;
        // The next hdcx_set_scene builds with the new options.
        tracer->builtIndices.clear();
        tracer->builtVertexCount = 0;
    });
}

int hdcx_get_bvh_info(hdcx_tracer* tracer, hdcx_bvh_info* out_info)
{
    return Guard([&] {
        if (!tracer || !out_info) throw std::invalid_argument("null argument");
        *out_info = tracer->geometry ? tracer->bvhInfo : hdcx_bvh_info{};
    });
}

int hdcx_read_emissive_table(hdcx_tracer* tracer, float* out, uint32_t count)
{
    return Guard([&] {
        if (!tracer || !out) throw std::invalid_argument("null argument");
        if (!tracer->geometry) throw Failure(HDCX_ERR_NO_GEOMETRY, "no scene");
        if (count != tracer->emissiveEntries) throw std::invalid_argument("count is not the scene's triangle count");
        tracer->Activate();
        Check(hipMemcpy(out, tracer->emissive.as<float>(), static_cast<std::size_t>(count) * 16, hipMemcpyDeviceToHost),
              "emissive readback");
    });
}

int hdcx_read_dome_importance(hdcx_tracer* tracer, uint32_t light, uint32_t* out_width, uint32_t* out_height,
                              float* out_marginal, float* out_conditional, float* out_pdf)
{
    return Guard([&] {
        if (!tracer || !out_width || !out_height) throw std::invalid_argument("null argument");
        if (!tracer->geometry) throw Failure(HDCX_ERR_NO_GEOMETRY, "no scene");
        if (light >= tracer->importanceLayout.size()) throw std::invalid_argument("light index out of range");
        const ImportanceDesc& d = tracer->importanceLayout[light];
        *out_width = d.width;
        *out_height = d.height;
        tracer->Activate();
        const float* pool = tracer->importancePool.as<float>();
        const std::size_t texels = static_cast<std::size_t>(d.width) * d.height;
        if (out_marginal && d.height)
            Check(hipMemcpy(out_marginal, pool + d.marginal, d.height * sizeof(float), hipMemcpyDeviceToHost), "readback");
        if (out_conditional && texels)
            Check(hipMemcpy(out_conditional, pool + d.conditional, texels * sizeof(float), hipMemcpyDeviceToHost), "readback");
        if (out_pdf && texels)
            Check(hipMemcpy(out_pdf, pool + d.pdf, texels * sizeof(float), hipMemcpyDeviceToHost), "readback");
    });
}

int hdcx_set_radiance_cache(hdcx_tracer* tracer, const hdcx_radiance_cache_config* config)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        if (!config || !config->enabled) {
            tracer->rcEnabled = false;
            return;
        }
        const hdcx_radiance_cache_config c = NormalizedRcConfig(*config);
        tracer->rcConfig = c;
        tracer->EnsureRc(c.seed);
        tracer->rcEnabled = true;
    });
}

int hdcx_rc_get_params(hdcx_tracer* tracer, float* out_params)
{
    return Guard([&] {
        if (!tracer || !out_params) throw std::invalid_argument("null argument");
        tracer->EnsureRc(tracer->rcInitialised ? tracer->rcSeed : 0);
        Check(hipMemcpy(out_params, tracer->rcMaster.as<void>(), kRcParams * sizeof(float), hipMemcpyDeviceToHost), "rc params");
    });
}

int hdcx_rc_set_params(hdcx_tracer* tracer, const float* params)
{
    return Guard([&] {
        if (!tracer || !params) throw std::invalid_argument("null argument");
        tracer->EnsureRc(tracer->rcInitialised ? tracer->rcSeed : 0);
        Check(hipMemcpy(tracer->rcMaster.as<void>(), params, kRcParams * sizeof(float), hipMemcpyHostToDevice), "rc params");
        Check(hipMemset(tracer->rcM.as<void>(), 0, tracer->rcM.bytes()), "rc params");
        Check(hipMemset(tracer->rcV.as<void>(), 0, tracer->rcV.bytes()), "rc params");
        tracer->rcStep = 0;
        Check(LaunchRcPublish(tracer->Rc(), tracer->stream), "rc publish");
        Check(hipStreamSynchronize(tracer->stream), "rc params");
    });
}

int hdcx_rc_infer(hdcx_tracer* tracer, const hdcx_rc_query* queries, uint32_t count, float* out)
{
    return Guard([&] {
        if (!tracer || (count && (!queries || !out))) throw std::invalid_argument("null argument");
        if (count == 0) return;
        tracer->EnsureRc(tracer->rcInitialised ? tracer->rcSeed : 0);
        tracer->rcQueries.Upload(queries, static_cast<std::size_t>(count) * sizeof(RcQuery), tracer->stream);
        tracer->rcOut.Allocate(static_cast<std::size_t>(count) * 3 * sizeof(float));
        Check(LaunchRcInfer(tracer->RcInferNet(), tracer->rcQueries.as<RcQuery>(), nullptr, count, tracer->rcOut.as<float>(),
                            nullptr, nullptr, tracer->maxBlocks, tracer->stream),
              "rc infer");
        Check(hipMemcpyAsync(out, tracer->rcOut.as<void>(), static_cast<std::size_t>(count) * 3 * sizeof(float),
                             hipMemcpyDeviceToHost, tracer->stream),
              "rc infer");
        Check(hipStreamSynchronize(tracer->stream), "rc infer");
    });
}

int hdcx_rc_gradient(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, float* out_grad, float* out_loss)
{
    return Guard([&] {
        if (!tracer || !samples || !out_grad || count == 0) throw std::invalid_argument("bad argument");
        tracer->EnsureRc(tracer->rcInitialised ? tracer->rcSeed : 0);
        hipStream_t st = tracer->stream;
        tracer->rcQueries.Upload(samples, static_cast<std::size_t>(count) * sizeof(RcQuery), st);
        tracer->rcOut.Allocate(std::max<std::size_t>(static_cast<std::size_t>(count) * 3, kRcParams) * sizeof(float));
        Check(hipMemsetAsync(tracer->rcLoss.as<void>(), 0, 2 * sizeof(float), st), "rc loss");
        tracer->rcAct.Allocate(std::max(tracer->rcAct.bytes(), RcGradWorkspaceBytes(count)));
        Check(LaunchRcGradient(tracer->RcNet(), tracer->rcQueries.as<RcQuery>(), count, tracer->rcAct.as<__bf16>(),
                               tracer->rcGrad.as<float>(), tracer->rcLoss.as<float>(), st),
              "rc gradient");
        Check(LaunchRcReduce(tracer->rcGrad.as<float>(), kRcGradSplits, tracer->rcOut.as<float>(), st), "rc reduce");
        float loss[2];
        Check(hipMemcpyAsync(out_grad, tracer->rcOut.as<void>(), kRcParams * sizeof(float), hipMemcpyDeviceToHost, st), "rc grad");
        Check(hipMemcpyAsync(loss, tracer->rcLoss.as<void>(), sizeof(loss), hipMemcpyDeviceToHost, st), "rc grad");
        Check(hipStreamSynchronize(st), "rc gradient");
        if (out_loss) *out_loss = loss[1] > 0.0F ? loss[0] / loss[1] : 0.0F;
    });
}

} // extern "C"

// ---- Scratch of hdcx_rc_gradient_noalloc ----
// The training workspace for up to kNoAllocMaxSamples samples, the
// partial gradient rows and the loss sums live in one file-scope device
// array (every GPU the library runs on holds its own copy for the life of
// the process). Host code never passes `scratch` itself to a kernel: its
// address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current.
namespace {
constexpr std::uint32_t kNoAllocMaxSamples = 4096;
constexpr std::size_t kNoAllocActBytes = hdcx::RcGradWorkspaceBytes(kNoAllocMaxSamples);
constexpr std::size_t kNoAllocPartialBytes = static_cast<std::size_t>(hdcx::kRcGradSplits) * hdcx::kRcParams * sizeof(float);
constexpr std::size_t kNoAllocScratchBytes = kNoAllocActBytes + kNoAllocPartialBytes + 256;  // + the loss sums
static_assert(kNoAllocActBytes % 256 == 0 && kNoAllocPartialBytes % 256 == 0, "aligned scratch regions");

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex gScratchMutex;
void* gScratchAddress[64] = {};

// Device address of `scratch` on the current device `device`.
unsigned char* ScratchAddress(int device)
{
    if (device < 0 || device >= 64) throw std::invalid_argument("device ordinal out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddress[device]) {
        void* ptr = nullptr;
        Check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch address");
        gScratchAddress[device] = ptr;
    }
    return static_cast<unsigned char*>(gScratchAddress[device]);
}
} // namespace

extern "C" {

int hdcx_rc_gradient_noalloc(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, float* out_grad,
                             float* out_loss)
{
    return Guard([&] {
        if (!tracer || !samples || !out_grad) throw std::invalid_argument("null argument");
        if (count == 0 || count > kNoAllocMaxSamples)
            throw std::invalid_argument("hdcx_rc_gradient_noalloc takes 1 to 4096 samples per call (its scratch size)");
        if (!tracer->rcInitialised)
            throw std::invalid_argument("the radiance cache has no network yet: call hdcx_set_radiance_cache first");
        tracer->Activate();
        unsigned char* base = ScratchAddress(tracer->device);
        auto* act = reinterpret_cast<__bf16*>(base);
        auto* partial = reinterpret_cast<float*>(base + kNoAllocActBytes);
        auto* loss = reinterpret_cast<float*>(base + kNoAllocActBytes + kNoAllocPartialBytes);
        hipStream_t st = tracer->stream;
        Check(LaunchRcClearLoss(loss, st), "rc loss");
        Check(LaunchRcGradient(tracer->RcNet(), reinterpret_cast<const RcQuery*>(samples), count, act, partial, loss, st),
              "rc gradient");
        Check(LaunchRcReduce(partial, kRcGradSplits, out_grad, st), "rc reduce");
        if (out_loss) Check(LaunchRcMeanLoss(loss, out_loss, st), "rc loss");
        Check(hipStreamSynchronize(st), "rc gradient");
    });
}

int hdcx_rc_train(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, uint32_t steps, float* out_loss)
{
    return Guard([&] {
        if (!tracer || !samples || count == 0) throw std::invalid_argument("bad argument");
        tracer->EnsureRc(tracer->rcInitialised ? tracer->rcSeed : 0);
        if (tracer->rcConfig.learning_rate <= 0.0F) {
            tracer->rcConfig.learning_rate = 1e-2F;
            tracer->rcConfig.beta1 = 0.9F;
            tracer->rcConfig.beta2 = 0.99F;
            tracer->rcConfig.epsilon = 1e-8F;
        }
        tracer->rcQueries.Upload(samples, static_cast<std::size_t>(count) * sizeof(RcQuery), tracer->stream);
        const float loss = tracer->RcTrain(tracer->rcQueries.as<RcQuery>(), nullptr, count, steps);
        if (out_loss) *out_loss = loss;
    });
}

int hdcx_rc_get_stats(hdcx_tracer* tracer, hdcx_rc_stats* out_stats)
{
    return Guard([&] {
        if (!tracer || !out_stats) throw std::invalid_argument("null argument");
        if (tracer->rcCounters.bytes()) {
            unsigned long long queries = 0;
            Check(hipMemcpy(&queries, tracer->rcCounters.as<std::uint32_t>() + 2, sizeof(queries), hipMemcpyDeviceToHost),
                  "rc stats");
            tracer->rcStats.queries = queries;
        }
        tracer->rcStats.calibration_scale = 1.0F;
        if (const float* scale = tracer->CalibrationScale())
            Check(hipMemcpy(&tracer->rcStats.calibration_scale, scale, sizeof(float), hipMemcpyDeviceToHost), "rc stats");
        *out_stats = tracer->rcStats;
    });
}

int hdcx_set_instrumentation(hdcx_tracer* tracer, uint32_t flags)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        tracer->instrumentation = flags;
    });
}

int hdcx_get_kernel_stats(hdcx_tracer* tracer, hdcx_kernel_stats* out_stats)
{
    return Guard([&] {
        if (!tracer || !out_stats) throw std::invalid_argument("null argument");
        tracer->Activate();
        hdcx_kernel_stats s = tracer->stats;
        if (tracer->totals.bytes()) {
            unsigned long long totals[5] = {0, 0, 0, 0, 0};
            TraversalCounters ctr[2] = {};
            Check(hipMemcpy(totals, tracer->totals.as<void>(), sizeof(totals), hipMemcpyDeviceToHost), "stats");
            Check(hipMemcpy(ctr, tracer->traversalCounters.as<void>(), sizeof(ctr), hipMemcpyDeviceToHost), "stats");
            s.rays[HDCX_KERNEL_EXTEND] = totals[3];
            s.rays[HDCX_KERNEL_MISS] = totals[0];
            s.rays[HDCX_KERNEL_SHADE] = totals[4];
            s.rays[HDCX_KERNEL_SHADOW] = totals[2];
            s.nodes_visited = ctr[0].nodes;
            s.triangles_tested = ctr[0].triangles;
            s.shadow_nodes_visited = ctr[1].nodes;
            s.shadow_triangles_tested = ctr[1].triangles;
            s.extend_rays_counted = (tracer->instrumentation & 2U) ? totals[3] : 0;
            s.extend_hits_counted = (tracer->instrumentation & 2U) ? ctr[0].hits : 0;
            s.shadow_queued_paths = totals[1];
            s.shadow_rays_counted = (tracer->instrumentation & 2U) ? totals[2] : 0;
        }
        *out_stats = s;
    });
}

int hdcx_reset_kernel_stats(hdcx_tracer* tracer)
{
    return Guard([&] {
        if (!tracer) throw std::invalid_argument("tracer is null");
        tracer->Activate();
        const std::uint64_t nodeBytes = tracer->stats.node_bytes, triBytes = tracer->stats.triangle_bytes;
        const std::uint64_t nodes = tracer->stats.bvh_nodes, tris = tracer->stats.bvh_triangles;
        tracer->stats = hdcx_kernel_stats{};
        tracer->stats.node_bytes = nodeBytes;
        tracer->stats.triangle_bytes = triBytes;
        tracer->stats.bvh_nodes = nodes;
        tracer->stats.bvh_triangles = tris;
        if (tracer->totals.bytes()) {
            Check(hipMemset(tracer->totals.as<void>(), 0, tracer->totals.bytes()), "reset");
            Check(hipMemset(tracer->traversalCounters.as<void>(), 0, tracer->traversalCounters.bytes()), "reset");
        }
    });
}

} // extern "C"
