// Batched multi-scene renderer: many independent scenes share one wavefront
// path pool; every stage is one launch over all scenes, and the per-wave
// launch sequence is captured once in a hipGraph and replayed for every
// wave (the wave index lives in device memory). Includes the autotuner of
// the wave / tile / grid sizes and one radiance cache per scene. C ABI in
// include/hdcx.h.
#include "../../include/hdcx.h"
#include "device_buffer.h"
#include "kernel_api.h"
#include "rc_config.h"
#include "rcache.h"
#include "scene_store.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using namespace hdcx;

namespace {

struct ViewState {
    hdcx_camera camera{};
    std::uint32_t width = 0, height = 0, samples = 0;
    bool set = false;
    bool zUp = false;
    DeviceBuffer mean, batch, out;
    std::uint32_t filmWidth = 0, filmHeight = 0;
};

// Tiles of every view packed into waves of at most `capacity` lane paths.
struct Plan {
    std::vector<BatchTile> tiles;
    std::vector<std::uint32_t> waveTiles{0U};
    std::vector<std::uint32_t> wavePaths, wavePixels;
    std::uint32_t maxPaths = 0, maxPixels = 0;
    std::uint64_t pixelSamples = 0, lanePaths = 0;
};

struct GraphKey {
    const void* pool = nullptr;
    const void* optional[3] = {};  // walk slots, path spread, scene index (allocated on demand)
    const void* tables[7] = {};
    std::uint32_t capacity = 0, bounces = 0, maxBlocks = 0, maxPaths = 0, maxPixels = 0, walk = 0;
    std::uint32_t kernels = 0;  // bit 0: two-level; bit 1: every scene opaque
    std::uint32_t scenes = 0;  // also keeps the struct free of implicit padding (keys compare with memcmp)
    CacheParams cache{};
    const void* rc[6] = {};    // per-scene weights, sorted queries, bucket arrays
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};
static_assert(sizeof(CacheParams) == 10 * 4 + 10 * 8, "CacheParams has implicit padding");
static_assert(sizeof(GraphKey) == 11 * 8 + 8 * 4 + sizeof(CacheParams) + 6 * 8, "GraphKey has implicit padding");

// Capacity-preserving upload (stable device addresses for a captured graph).
void Put(DeviceBuffer& b, const void* host, std::size_t bytes, hipStream_t st)
{
    b.Reserve(std::max<std::size_t>(bytes, 16));
    if (bytes) Check(hipMemcpyAsync(b.as<void>(), host, bytes, hipMemcpyHostToDevice, st), "batch upload");
}

} // namespace

struct hdcx_batch : PathPool {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<SceneStore>> scenes;
    std::vector<ViewState> views;
    hdcx_batch_tuning tuning{3U << 20, 4096U, 2048U, 1U};
    hdcx_batch_stats stats{};
    // Per-render tables (host copies stay alive until the render syncs).
    Plan plan;
    std::vector<BatchView> viewHost;
    std::vector<DeviceScene> sceneHost;
    DeviceBuffer sceneTable, viewTable, tiles, waveTiles, wavePaths, wavePixels, cursor;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    GraphKey graphKey{};

    // Radiance caches: the networks of all scenes side by side (scene s at
    // s * kRcParams; transposes at s * 4 * 4096), trained after each render
    // on the samples it gathered (sampleCapacity per scene).
    static constexpr std::uint32_t kRecordVertices = 4;
    hdcx_radiance_cache_config rcConfig{};
    bool rcEnabled = false;
    std::uint64_t rcSeed = 0;
    std::uint32_t rcScenes = 0, rcFrame = 0;
    std::vector<int> rcStep;
    DeviceBuffer rcMaster, rcM, rcV, rcEma, rcW, rcWt, rcWe, rcWte, rcGrad, rcLoss, rcAct;
    DeviceBuffer rcQueries, rcSorted, rcCounters, rcRecords, rcRecordThroughput, rcTrainSlot, rcSamples, rcSceneSamples,
        rcSceneTrain, rcBucketCount, rcBucketStart, rcBucketBlocks, rcBucketCursor;
    hdcx_batch_rc_stats rcStats{};
    std::vector<uint2> rcTrainHost;  // host copy of the per-scene training table (alive until the render syncs)

    ~hdcx_batch()
    {
        DropGraph();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void Activate() const { Check(hipSetDevice(device), "hipSetDevice"); }
    void DropGraph()
    {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
        graphKey = GraphKey{};
    }

    RcState RcSceneState(std::uint32_t k)
    {
        const std::size_t p = static_cast<std::size_t>(k) * kRcParams, t = static_cast<std::size_t>(k) * 4 * 4096;
        RcState r{};
        r.master = rcMaster.as<float>() + p;
        r.m = rcM.as<float>() + p;
        r.v = rcV.as<float>() + p;
        r.grad = rcGrad.as<float>();
        r.w = rcW.as<__bf16>() + p;
        r.wt = rcWt.as<__bf16>() + t;
        r.ema = rcEma.as<float>() + p;
        r.we = rcWe.as<__bf16>() + p;
        r.wte = rcWte.as<__bf16>() + t;
        r.emaDecay = rcConfig.ema_decay;
        r.loss = rcLoss.as<float>();
        return r;
    }

    // Weights of every scene from seed + scene index, Adam state cleared.
    void EnsureRc()
    {
        const auto n = static_cast<std::uint32_t>(scenes.size());
        if (rcScenes == n && rcSeed == rcConfig.seed && rcMaster.bytes()) return;
        const std::size_t params = static_cast<std::size_t>(n) * kRcParams;
        rcMaster.Allocate(params * 4);
        rcM.Allocate(params * 4);
        rcV.Allocate(params * 4);
        rcEma.Allocate(params * 4);
        rcW.Allocate(params * 2);
        rcWe.Allocate(params * 2);
        rcWt.Allocate(static_cast<std::size_t>(n) * 4 * 4096 * 2);
        rcWte.Allocate(static_cast<std::size_t>(n) * 4 * 4096 * 2);
        rcGrad.Allocate(static_cast<std::size_t>(kRcGradSplits) * kRcParams * 4);
        rcLoss.Allocate(8);
        std::vector<float> w(params);
        for (std::uint32_t k = 0; k < n; ++k) InitRcWeights(rcConfig.seed + k, w.data() + static_cast<std::size_t>(k) * kRcParams);
        Check(hipMemcpyAsync(rcMaster.as<void>(), w.data(), params * 4, hipMemcpyHostToDevice, stream), "batch rc init");
        Check(hipMemsetAsync(rcM.as<void>(), 0, rcM.bytes(), stream), "batch rc init");
        Check(hipMemsetAsync(rcV.as<void>(), 0, rcV.bytes(), stream), "batch rc init");
        Check(hipMemsetAsync(rcGrad.as<void>(), 0, rcGrad.bytes(), stream), "batch rc init");
        for (std::uint32_t k = 0; k < n; ++k) Check(LaunchRcPublish(RcSceneState(k), stream), "batch rc publish");
        Check(hipStreamSynchronize(stream), "batch rc init");
        rcStep.assign(n, 0);
        rcScenes = n;
        rcSeed = rcConfig.seed;
        rcFrame = 0;
    }

    // Cache hooks of one render: per-scene training strides (about
    // train_samples samples per scene: 3 lanes x ~2.5 recorded vertices per
    // training pixel sample), record slots for the whole frame, query queue
    // for a wave.
    CacheParams PrepareRc(std::uint32_t maxPaths)
    {
        EnsureRc();
        const auto n = static_cast<std::uint32_t>(scenes.size());
        std::vector<uint2>& train = rcTrainHost;
        train.assign(n, uint2{0U, 0U});
        std::uint64_t slots = 0;
        const std::uint64_t wanted = std::max<std::uint64_t>(1, rcConfig.train_samples * 2ULL / 15ULL);
        for (std::uint32_t k = 0; k < n; ++k) {
            const ViewState& v = views[k];
            if (!v.set || v.samples == 0 || !scenes[k]->geometry) continue;
            const std::uint64_t frame = static_cast<std::uint64_t>(v.width) * v.height * v.samples;
            const auto stride = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, frame / wanted));
            train[k] = uint2{stride, static_cast<std::uint32_t>((rcFrame * 2654435761ULL) % stride)};
            slots += (frame / stride + 2) * 3;
        }
        const std::uint32_t cap = std::max(rcConfig.train_samples, 1U);
        const std::size_t records = static_cast<std::size_t>(slots) * kRecordVertices;
        if (rcRecords.bytes() < records * sizeof(RcQuery)) {
            rcRecords.Allocate(records * sizeof(RcQuery));
            // Every record starts unused (lane ~0); the gather marks consumed ones so again.
            Check(hipMemsetAsync(rcRecords.as<void>(), 0xff, rcRecords.bytes(), stream), "batch rc records");
            rcRecordThroughput.Allocate(records * 4);
        }
        rcQueries.Reserve(static_cast<std::size_t>(maxPaths) * sizeof(RcQuery));
        rcSorted.Reserve(static_cast<std::size_t>(maxPaths) * sizeof(RcQuery));
        rcTrainSlot.Reserve(static_cast<std::size_t>(maxPaths) * 4);
        rcSamples.Reserve(static_cast<std::size_t>(n) * cap * sizeof(RcQuery));
        rcSceneSamples.Reserve(static_cast<std::size_t>(n) * 4);
        rcBucketCount.Reserve(static_cast<std::size_t>(n) * 4);
        rcBucketCursor.Reserve(static_cast<std::size_t>(n) * 4);
        rcBucketStart.Reserve(static_cast<std::size_t>(n + 1) * 4);
        rcBucketBlocks.Reserve(static_cast<std::size_t>(n + 1) * 4);
        rcCounters.Reserve(64);
        Check(hipMemsetAsync(rcCounters.as<void>(), 0, 64, stream), "batch rc counters");
        Check(hipMemsetAsync(rcSceneSamples.as<void>(), 0, static_cast<std::size_t>(n) * 4, stream), "batch rc counters");
        Check(hipMemsetAsync(rcBucketCount.as<void>(), 0, static_cast<std::size_t>(n) * 4, stream), "batch rc counters");
        Put(rcSceneTrain, train.data(), train.size() * sizeof(uint2), stream);

        CacheParams cp{};
        cp.enabled = 1;
        cp.terminateBounce = rcConfig.terminate_bounce;
        cp.spreadThreshold = rcConfig.spread_threshold;
        cp.trainStride = 1;
        cp.recordVertices = kRecordVertices;
        cp.queryCapacity = maxPaths;
        cp.recordSlots = static_cast<std::uint32_t>(slots);
        cp.sampleCapacity = cap;
        auto* counters = rcCounters.as<std::uint32_t>();
        cp.queries = rcQueries.as<RcQuery>();
        cp.queryCount = counters;        // [0]; [2..3]: queries answered (64-bit)
        cp.recordCount = counters + 4;
        cp.records = rcRecords.as<RcQuery>();
        cp.recordThroughput = rcRecordThroughput.as<float>();
        cp.samples = rcSamples.as<RcQuery>();
        cp.sampleCount = counters + 1;
        cp.trainSlot = rcTrainSlot.as<std::uint32_t>();
        cp.sceneTrain = rcSceneTrain.as<uint2>();
        cp.sceneSampleCount = rcSceneSamples.as<std::uint32_t>();
        return cp;
    }

    RcSceneBuckets Buckets()
    {
        RcSceneBuckets b{};
        b.count = rcBucketCount.as<std::uint32_t>();
        b.start = rcBucketStart.as<std::uint32_t>();
        b.blockStart = rcBucketBlocks.as<std::uint32_t>();
        b.cursor = rcBucketCursor.as<std::uint32_t>();
        b.scenes = static_cast<std::uint32_t>(scenes.size());
        return b;
    }

    // Adam steps per scene on the samples the render gathered.
    void TrainRc()
    {
        const auto t0 = std::chrono::steady_clock::now();
        const auto n = static_cast<std::uint32_t>(scenes.size());
        std::vector<std::uint32_t> counts(n, 0U);
        std::uint64_t answered = 0;
        Check(hipMemcpyAsync(counts.data(), rcSceneSamples.as<void>(), n * 4, hipMemcpyDeviceToHost, stream), "batch rc counts");
        Check(hipMemcpyAsync(&answered, rcCounters.as<std::uint32_t>() + 2, 8, hipMemcpyDeviceToHost, stream), "batch rc counts");
        Check(hipStreamSynchronize(stream), "batch rc counts");
        Check(hipMemsetAsync(rcLoss.as<void>(), 0, 8, stream), "batch rc loss");
        const std::uint32_t cap = std::max(rcConfig.train_samples, 1U);
        const std::uint32_t steps = std::max(rcConfig.train_steps, 1U);
        rcAct.Allocate(std::max(rcAct.bytes(), RcGradWorkspaceBytes(cap / steps + 1U)));
        std::uint32_t trained = 0;
        for (std::uint32_t k = 0; k < n; ++k) {
            const std::uint32_t count = std::min(counts[k], cap);
            if (count == 0) continue;
            ++trained;
            const RcQuery* samples = rcSamples.as<RcQuery>() + static_cast<std::size_t>(k) * cap;
            const RcState st = RcSceneState(k);
            for (std::uint32_t i = 0; i < steps; ++i) {
                const auto b0 = static_cast<std::uint32_t>(static_cast<std::uint64_t>(count) * i / steps);
                const auto b1 = static_cast<std::uint32_t>(static_cast<std::uint64_t>(count) * (i + 1) / steps);
                if (b1 <= b0) continue;
                Check(LaunchRcGradient(RcNetwork{st.w, st.wt}, samples + b0, b1 - b0, rcAct.as<__bf16>(), st.grad, st.loss,
                                       stream),
                      "batch rc gradient");
                Check(LaunchRcAdamBlocks(st, kRcGradSplits, ++rcStep[k], rcConfig.learning_rate,
                                         rcConfig.beta1, rcConfig.beta2, rcConfig.epsilon, stream),
                      "batch rc adam");
                ++rcStats.adam_steps;
                rcStats.train_samples += b1 - b0;
            }
        }
        float loss[2] = {0.0F, 0.0F};
        Check(hipMemcpyAsync(loss, rcLoss.as<void>(), 8, hipMemcpyDeviceToHost, stream), "batch rc loss");
        Check(hipStreamSynchronize(stream), "batch rc train");
        rcStats.queries = answered;
        rcStats.mean_loss = loss[1] > 0.0F ? loss[0] / loss[1] : 0.0F;
        rcStats.scenes_trained = trained;
        rcStats.train_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ++rcFrame;
    }

    void MakePlan()
    {
        plan = Plan{};
        const std::uint32_t cap = std::max(tuning.wave_paths, 3U);
        std::vector<BatchTile> all;
        for (std::uint32_t s = 0; s < views.size(); ++s) {
            const ViewState& v = views[s];
            if (!v.set || v.samples == 0 || !scenes[s]->geometry) continue;
            const std::uint64_t per = 3ULL * v.samples;
            if (per > cap) throw std::invalid_argument("wave_paths is below 3 x samples of a view");
            const auto tile = static_cast<std::uint32_t>(
                std::max<std::uint64_t>(1, std::min<std::uint64_t>(std::max(tuning.tile_pixels, 1U), cap / per)));
            const std::uint32_t pixels = v.width * v.height;
            for (std::uint32_t p0 = 0; p0 < pixels; p0 += tile) {
                BatchTile t{};
                t.scene = s;
                t.p0 = p0;
                t.p1 = std::min(pixels, p0 + tile);
                all.push_back(t);
            }
            plan.pixelSamples += static_cast<std::uint64_t>(pixels) * v.samples;
        }
        std::uint32_t paths = 0, pixels = 0;
        for (BatchTile t : all) {
            const std::uint32_t tp = (t.p1 - t.p0) * views[t.scene].samples * 3U;
            if (paths + tp > cap) {
                plan.waveTiles.push_back(static_cast<std::uint32_t>(plan.tiles.size()));
                plan.wavePaths.push_back(paths);
                plan.wavePixels.push_back(pixels);
                paths = pixels = 0;
            }
            t.pathOffset = paths;
            t.pixelOffset = pixels;
            paths += tp;
            pixels += t.p1 - t.p0;
            plan.tiles.push_back(t);
            plan.lanePaths += tp;
        }
        if (paths > 0) {
            plan.waveTiles.push_back(static_cast<std::uint32_t>(plan.tiles.size()));
            plan.wavePaths.push_back(paths);
            plan.wavePixels.push_back(pixels);
        }
        for (std::uint32_t p : plan.wavePaths) plan.maxPaths = std::max(plan.maxPaths, p);
        for (std::uint32_t p : plan.wavePixels) plan.maxPixels = std::max(plan.maxPixels, p);
    }

    void Render(std::uint32_t sampleIndex, std::uint32_t maxBounces)
    {
        Activate();
        const auto t0 = std::chrono::steady_clock::now();
        const std::uint32_t bounces = std::clamp(maxBounces, 1U, 12U);
        const unsigned maxBlocks = std::max(tuning.max_blocks, 1U);
        MakePlan();
        stats = hdcx_batch_stats{};
        if (plan.tiles.empty()) return;
        hipStream_t st = stream;

        viewHost.assign(views.size(), BatchView{});
        sceneHost.assign(scenes.size(), DeviceScene{});
        for (std::uint32_t s = 0; s < views.size(); ++s) {
            ViewState& v = views[s];
            sceneHost[s] = scenes[s]->scene;
            if (!v.set) continue;
            sceneHost[s].zUp = v.zUp ? 1U : 0U;
            const std::size_t n = static_cast<std::size_t>(v.width) * v.height;
            if (v.filmWidth != v.width || v.filmHeight != v.height || !v.mean.bytes()) {
                v.mean.Allocate(n * sizeof(double4));
                v.batch.Allocate(n * sizeof(double4));
                v.out.Allocate(n * sizeof(float4));
                v.filmWidth = v.width;
                v.filmHeight = v.height;
                Check(LaunchZero(v.mean.as<double4>(), static_cast<std::uint32_t>(n), maxBlocks, st), "zero film");
            }
            Check(LaunchZero(v.batch.as<double4>(), static_cast<std::uint32_t>(n), maxBlocks, st), "zero batch");
            BatchView& bv = viewHost[s];
            std::memcpy(bv.cam + 0, v.camera.origin, 12);
            for (int k = 0; k < 3; ++k)
                bv.cam[3 + k] = static_cast<float>(static_cast<double>(v.camera.lower_left[k]) -
                                                   static_cast<double>(v.camera.origin[k]));
            std::memcpy(bv.cam + 6, v.camera.horizontal, 12);
            std::memcpy(bv.cam + 9, v.camera.vertical, 12);
            bv.width = v.width;
            bv.height = v.height;
            bv.sampleBase = sampleIndex;
            bv.samples = v.samples;
            bv.batch = v.batch.as<double4>();
        }
        Put(sceneTable, sceneHost.data(), sceneHost.size() * sizeof(DeviceScene), st);
        Put(viewTable, viewHost.data(), viewHost.size() * sizeof(BatchView), st);
        Put(tiles, plan.tiles.data(), plan.tiles.size() * sizeof(BatchTile), st);
        Put(waveTiles, plan.waveTiles.data(), plan.waveTiles.size() * 4, st);
        Put(wavePaths, plan.wavePaths.data(), plan.wavePaths.size() * 4, st);
        Put(wavePixels, plan.wavePixels.data(), plan.wavePixels.size() * 4, st);
        const std::uint32_t zero = 0U;
        Put(cursor, &zero, 4, st);

        BatchWaves bw{};
        bw.views = viewTable.as<BatchView>();
        bw.tiles = tiles.as<BatchTile>();
        bw.waveTiles = waveTiles.as<std::uint32_t>();
        bw.wavePaths = wavePaths.as<std::uint32_t>();
        bw.wavePixels = wavePixels.as<std::uint32_t>();
        bw.cursor = cursor.as<std::uint32_t>();
        const DeviceScene* table = sceneTable.as<DeviceScene>();
        bool twoLevel = false, subsurface = false, simple = true, opaque = true;
        for (const DeviceScene& d : sceneHost) {
            twoLevel = twoLevel || d.instanceCount > 0;
            opaque = opaque && d.opaque;
            subsurface = subsurface || d.hasSubsurface;
            simple = simple && d.simpleMaterials;
        }
        const int walk = simple ? kShadeSimple : subsurface ? kShadeWalkDefer : kShadeWalkInline;
        EnsurePaths(plan.maxPaths, kNeedSceneIndex | (walk == kShadeWalkDefer ? kNeedWalks : 0U) |
                                       (rcEnabled ? kNeedSpread : 0U));
        const PathBuffers pb = Paths();
        const std::uint32_t maxPaths = plan.maxPaths, maxPixels = plan.maxPixels;
        const bool cache = rcEnabled;
        const CacheParams cp = cache ? PrepareRc(maxPaths) : CacheParams{};
        const RcSceneBuckets buckets = cache ? Buckets() : RcSceneBuckets{};
        auto enqueueWave = [&] {
            Check(LaunchRaygenBatch(pb, bw, cp, maxPaths, maxBlocks, st), "batch raygen");
            for (std::uint32_t b = 0; b < bounces; ++b) {
                const std::uint32_t in = b & 1U;
                Check(LaunchExtendBatch(table, twoLevel, opaque, pb, in, b == 0 ? kShareCamera : kShareLanes, maxPaths,
                                        maxBlocks, nullptr, st),
                      "batch extend");
                const int last = b + 1 == bounces ? 1 : 0;
                Check(b == 0 ? LaunchMissCameraBatch(table, pb, maxPaths, maxBlocks, st)
                             : LaunchMissBatch(table, pb, in, maxPaths, maxBlocks, st),
                      "batch miss");
                // Shade also pairs the lanes of each pixel sample with equal
                // next rays and equal shadow rays.
                Check(LaunchShadeBatch(table, twoLevel, pb, in, static_cast<int>(b), last, cp, walk, 1, maxPaths, maxBlocks,
                                       st),
                      "batch shade");
                if (walk == kShadeWalkDefer && !last) {
                    Check(LaunchWalkBatch(table, twoLevel, pb, maxPaths, maxBlocks, st), "batch walk");
                    Check(LaunchShadeBatch(table, twoLevel, pb, in, static_cast<int>(b), last, cp, kShadeWalkResume, 1,
                                           maxPaths, maxBlocks, st),
                          "batch shade (resume)");
                }
                if (cache && (b == cp.terminateBounce || (b > cp.terminateBounce && cp.spreadThreshold > 0.0F))) {
                    // Queries of this bounce, answered by their scenes' networks.
                    Check(LaunchRcInferScenes(rcWe.as<__bf16>(), cp.queries, cp.queryCount, maxPaths, pb.pathScene, buckets,
                                              rcSorted.as<RcQuery>(), pb.radiance, pb.throughput, maxBlocks, st),
                          "batch rc infer");
                    Check(LaunchConsumeQueries(cp, st), "batch rc queries");
                }
                Check(LaunchShadowBatch(table, twoLevel, opaque, pb, maxPaths, maxBlocks, nullptr, st), "batch shadow");
            }
            if (cache)
                Check(LaunchGatherTraining(pb, cp, cp.recordSlots * kRecordVertices, maxBlocks, st), "batch rc samples");
            Check(LaunchFilmBatch(pb, bw, maxPixels, maxBlocks, st), "batch film");
            Check(LaunchAdvanceWave(bw, cache ? cp.recordCount : nullptr, st), "batch cursor");
        };
        const auto waves = static_cast<std::uint32_t>(plan.wavePaths.size());
        if (tuning.use_graph) {
            GraphKey key{};
            key.pool = rayO.as<void>();
            key.optional[0] = this->walk.as<void>();
            key.optional[1] = spread.as<void>();
            key.optional[2] = pathScene.as<void>();
            const void* t[7] = {table, bw.views, bw.tiles, bw.waveTiles, bw.wavePaths, bw.wavePixels, bw.cursor};
            std::memcpy(key.tables, t, sizeof(t));
            key.capacity = capacity;
            key.bounces = bounces;
            key.maxBlocks = maxBlocks;
            key.maxPaths = maxPaths;
            key.maxPixels = maxPixels;
            key.kernels = (twoLevel ? 1U : 0U) | (opaque ? 2U : 0U);
            key.walk = static_cast<std::uint32_t>(walk);
            key.scenes = static_cast<std::uint32_t>(scenes.size());
            key.cache = cp;
            if (cache) {
                const void* r[6] = {rcWe.as<void>(), rcSorted.as<void>(), buckets.count, buckets.start, buckets.blockStart,
                                    buckets.cursor};
                std::memcpy(key.rc, r, sizeof(r));
            }
            if (!graphExec || !(key == graphKey)) {
                DropGraph();
                Check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
                try {
                    enqueueWave();
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                Check(hipStreamEndCapture(st, &graph), "graph capture");
                Check(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0), "graph instantiate");
                graphKey = key;
            }
            for (std::uint32_t w = 0; w < waves; ++w) Check(hipGraphLaunch(graphExec, st), "graph launch");
            stats.graph_launches = waves;
        } else {
            for (std::uint32_t w = 0; w < waves; ++w) enqueueWave();
        }
        for (std::uint32_t s = 0; s < views.size(); ++s) {
            ViewState& v = views[s];
            if (!v.set || v.samples == 0 || !scenes[s]->geometry) continue;
            Check(LaunchResolve(v.mean.as<double4>(), v.batch.as<double4>(), v.out.as<float4>(), v.width * v.height,
                                sampleIndex, v.samples, maxBlocks, st),
                  "batch resolve");
        }
        Check(hipStreamSynchronize(st), "batch render");
        if (cache) TrainRc();
        stats.pixel_samples = plan.pixelSamples;
        stats.lane_paths = plan.lanePaths;
        stats.waves = waves;
        stats.milliseconds = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    std::string TuningKey(std::uint32_t bounces) const
    {
        hipDeviceProp_t prop{};
        Check(hipGetDeviceProperties(&prop, device), "device properties");
        std::uint64_t pixels = 0, samples = 0;
        std::uint32_t active = 0;
        for (const ViewState& v : views) {
            if (!v.set) continue;
            ++active;
            pixels += static_cast<std::uint64_t>(v.width) * v.height;
            samples += static_cast<std::uint64_t>(v.width) * v.height * v.samples;
        }
        std::ostringstream k;
        k << prop.gcnArchName << "/cu" << prop.multiProcessorCount << "/scenes" << active << "/px" << pixels << "/ps"
          << samples << "/b" << bounces;
        return k.str();
    }

    hdcx_batch_tuning Autotune(std::uint32_t bounces, const char* cachePath)
    {
        const std::string key = TuningKey(bounces);
        if (cachePath) {
            std::ifstream in(cachePath);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ls(line);
                std::string k;
                hdcx_batch_tuning t{};
                if (ls >> k >> t.wave_paths >> t.tile_pixels >> t.max_blocks >> t.use_graph && k == key) {
                    tuning = t;
                    return tuning;
                }
            }
        }
        std::vector<std::uint32_t> saved;
        for (ViewState& v : views) {
            saved.push_back(v.samples);
            v.samples = std::min(v.samples, 1U);
        }
        auto measure = [&](const hdcx_batch_tuning& t) {
            tuning = t;
            Render(0, bounces);  // warm-up (allocations, graph capture)
            double best = 1e300;
            for (int r = 0; r < 2; ++r) {
                Render(0, bounces);
                best = std::min(best, stats.milliseconds);
            }
            return best;
        };
        hdcx_batch_tuning best{3U << 20, 4096U, 2048U, 1U};
        double bestMs = 1e300;
        try {
            bestMs = measure(best);
            // Coordinate descent: grid size, then wave size, then tile size.
            for (std::uint32_t mb : {1024U, 4096U}) {
                hdcx_batch_tuning t = best;
                t.max_blocks = mb;
                const double ms = measure(t);
                if (ms < bestMs) { bestMs = ms; best = t; }
            }
            for (std::uint32_t wp : {3U << 19, 3U << 21, 3U << 23}) {
                hdcx_batch_tuning t = best;
                t.wave_paths = wp;
                const double ms = measure(t);
                if (ms < bestMs) { bestMs = ms; best = t; }
            }
            for (std::uint32_t tp : {1024U, 16384U}) {
                hdcx_batch_tuning t = best;
                t.tile_pixels = tp;
                const double ms = measure(t);
                if (ms < bestMs) { bestMs = ms; best = t; }
            }
        } catch (...) {
            for (std::size_t i = 0; i < views.size(); ++i) views[i].samples = saved[i];
            throw;
        }
        for (std::size_t i = 0; i < views.size(); ++i) views[i].samples = saved[i];
        tuning = best;
        if (cachePath) {
            std::ofstream out(cachePath, std::ios::app);
            out << key << ' ' << best.wave_paths << ' ' << best.tile_pixels << ' ' << best.max_blocks << ' '
                << best.use_graph << '\n';
        }
        return tuning;
    }
};

extern "C" {

int hdcx_batch_create(int device_ordinal, hdcx_batch** out_batch)
{
    return Guard([&] {
        if (!out_batch) throw std::invalid_argument("out_batch is null");
        *out_batch = nullptr;
        int count = 0;
        Check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device_ordinal < 0 || device_ordinal >= count) throw Failure(HDCX_ERR_DEVICE, "no such HIP device");
        auto b = std::make_unique<hdcx_batch>();
        b->device = device_ordinal;
        b->Activate();
        Check(hipStreamCreate(&b->stream), "hipStreamCreate");
        *out_batch = b.release();
    });
}

int hdcx_batch_destroy(hdcx_batch* batch)
{
    return Guard([&] {
        if (!batch) return;
        batch->Activate();
        (void)hipStreamSynchronize(batch->stream);
        delete batch;
    });
}

int hdcx_batch_add_scene(hdcx_batch* batch, const hdcx_scene_view* scene, const hdcx_bvh_options* options,
                         uint32_t* out_index)
{
    return Guard([&] {
        if (!batch || !scene || !out_index) throw std::invalid_argument("null argument");
        batch->Activate();
        auto store = std::make_unique<SceneStore>();
        if (options) store->bvhOptions = *options;
        UploadScene(store.get(), scene, batch->stream);
        if (!store->geometry) throw Failure(HDCX_ERR_NO_GEOMETRY, "scene has no geometry");
        batch->scenes.push_back(std::move(store));
        batch->views.emplace_back();
        *out_index = static_cast<std::uint32_t>(batch->scenes.size() - 1);
    });
}

int hdcx_batch_set_view(hdcx_batch* batch, uint32_t scene, const hdcx_camera* camera, uint32_t width, uint32_t height,
                        uint32_t samples)
{
    return Guard([&] {
        if (!batch || !camera) throw std::invalid_argument("null argument");
        if (scene >= batch->views.size()) throw std::invalid_argument("scene index out of range");
        if (static_cast<std::uint64_t>(width) * height > (1ULL << 26)) throw std::invalid_argument("image too large");
        ViewState& v = batch->views[scene];
        v.camera = *camera;
        v.width = width;
        v.height = height;
        v.samples = samples;
        v.set = width > 0 && height > 0;
        // Fallback sky axis from the camera (vulkan_path_tracer.cpp:3237-3241).
        v.zUp = std::fabs(camera->vertical[2]) > std::fabs(camera->vertical[1]);
    });
}

int hdcx_batch_set_tuning(hdcx_batch* batch, const hdcx_batch_tuning* tuning)
{
    return Guard([&] {
        if (!batch || !tuning) throw std::invalid_argument("null argument");
        if (tuning->wave_paths < 3 || tuning->tile_pixels == 0 || tuning->max_blocks == 0)
            throw std::invalid_argument("tuning values must be positive (wave_paths >= 3)");
        batch->tuning = *tuning;
    });
}

int hdcx_batch_get_tuning(hdcx_batch* batch, hdcx_batch_tuning* out_tuning)
{
    return Guard([&] {
        if (!batch || !out_tuning) throw std::invalid_argument("null argument");
        *out_tuning = batch->tuning;
    });
}

int hdcx_batch_autotune(hdcx_batch* batch, uint32_t max_bounces, const char* cache_path, hdcx_batch_tuning* out_tuning)
{
    return Guard([&] {
        if (!batch) throw std::invalid_argument("batch is null");
        const hdcx_batch_tuning t = batch->Autotune(max_bounces, cache_path);
        if (out_tuning) *out_tuning = t;
    });
}

int hdcx_batch_render(hdcx_batch* batch, uint32_t sample_index, uint32_t max_bounces)
{
    return Guard([&] {
        if (!batch) throw std::invalid_argument("batch is null");
        batch->Render(sample_index, max_bounces);
    });
}

int hdcx_batch_read_film(hdcx_batch* batch, uint32_t scene, float* out_rgba)
{
    return Guard([&] {
        if (!batch || !out_rgba) throw std::invalid_argument("null argument");
        if (scene >= batch->views.size()) throw std::invalid_argument("scene index out of range");
        const ViewState& v = batch->views[scene];
        if (!v.out.bytes()) throw Failure(HDCX_ERR_NO_GEOMETRY, "nothing rendered for this scene");
        batch->Activate();
        Check(hipMemcpy(out_rgba, v.out.as<void>(), static_cast<std::size_t>(v.filmWidth) * v.filmHeight * 16,
                        hipMemcpyDeviceToHost),
              "film readback");
    });
}

int hdcx_batch_get_stats(hdcx_batch* batch, hdcx_batch_stats* out_stats)
{
    return Guard([&] {
        if (!batch || !out_stats) throw std::invalid_argument("null argument");
        *out_stats = batch->stats;
    });
}

int hdcx_batch_set_radiance_cache(hdcx_batch* batch, const hdcx_radiance_cache_config* config)
{
    return Guard([&] {
        if (!batch) throw std::invalid_argument("batch is null");
        if (!config || !config->enabled) {
            batch->rcEnabled = false;
            return;
        }
        batch->Activate();
        batch->rcConfig = NormalizedRcConfig(*config);
        batch->EnsureRc();
        batch->rcEnabled = true;
    });
}

int hdcx_batch_get_rc_stats(hdcx_batch* batch, hdcx_batch_rc_stats* out_stats)
{
    return Guard([&] {
        if (!batch || !out_stats) throw std::invalid_argument("null argument");
        *out_stats = batch->rcStats;
    });
}

} // extern "C"
