// Online-trained MLP radiance cache (no counterpart in the reference; its
// neural-rendering plan is unimplemented, docs/neural-rendering-plan.md:32-48).
//
// Network: 128 inputs (frequency-encoded position and direction, hero
// wavelength, roughness, albedo, a constant) -> 5 hidden layers of 64 ReLU
// units -> 3 outputs, the radiance scattered from a surface point toward a
// direction at the three wavelengths of a hero triple. No biases besides the
// constant input. BF16 weights/activations, FP32 accumulation on MFMA.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hdcx {

constexpr int kRcIn = 128;
constexpr int kRcHidden = 64;
constexpr int kRcOut = 3;
// Row-major weight offsets (elements): W0 [64][128], W1..W4 [64][64], W5 [3][64].
constexpr int kRcW0 = 0;
constexpr int kRcW1 = 8192;
constexpr int kRcW5 = 24576;
constexpr int kRcParams = 24768;
constexpr int kRcPositionBands = 12;
constexpr int kRcDirectionBands = 6;

// One cache query (48 bytes): position (normalised to the scene box),
// outgoing direction, hero-triple base wavelength, roughness, albedo, the
// lane (0..2) whose output is used, and the path it belongs to.
struct RcQuery {
    float px, py, pz;
    float dx, dy, dz;
    float lambda0;
    float roughness;
    float albedo;
    std::uint32_t lane;
    std::uint32_t path;
    float target;  // training samples: the lane's radiance estimate
};

struct RcNetwork {
    const __bf16* w;   // [kRcParams] row-major per layer
    const __bf16* wt;  // transposes of W1..W4 ([64][64] each) at kRcW1 + 4096 (l - 1)
};

struct RcState {
    float* master;     // FP32 weights [kRcParams]
    float* m;          // Adam first moment
    float* v;          // Adam second moment
    float* grad;       // accumulated gradient
    __bf16* w;         // BF16 weights (training passes)
    __bf16* wt;        // BF16 transposes of W1..W4
    float* loss;       // [2]: summed relative L2 loss, sample count
    float* ema;        // FP32 exponential moving average of the weights (from the published weights)
    __bf16* we;        // BF16 inference weights: the average (or the weights when emaDecay == 0)
    __bf16* wte;       // their transposes
    float emaDecay;    // per Adam step
};

// Inference: out[i * 3 + k] for every query (when out != nullptr), and
// radiance[path] += throughput[path] * scale * max(out[lane], 0) when
// radiance != nullptr, with scale = *calibration (1 when it is null).
hipError_t LaunchRcInfer(const RcNetwork& net, const RcQuery* queries, const std::uint32_t* count,
                         std::uint32_t maxCount, float* out, float* radiance, const float* throughput,
                         unsigned maxBlocks, hipStream_t st, const float* calibration = nullptr);

// Brightness calibration of the cache against the plain estimator it learns
// from. For the frame's training samples (targets t_i on lane k_i) and the
// network's outputs y on them, sums[0] += sum t_i and sums[1] += sum
// max(y_{k_i}, 0) (FP64). Then, in one thread: acc = decay * acc + sums,
// sums cleared, and *scale = acc[0] / acc[1] (1 while acc[1] is 0).
// Layout of `state`: double sums[2], double acc[2], float scale.
struct RcCalibration {
    double sums[2];
    double acc[2];
    float scale;
    float pad[3];
};
hipError_t LaunchRcCalibrate(const RcQuery* samples, std::uint32_t count, const float* outputs, RcCalibration* state,
                             float decay, hipStream_t st);
// Buckets of cache queries by scene (batch renderer), device arrays:
// count [scenes] (zero before the first use; the scan clears it), start and
// blockStart [scenes + 1], cursor [scenes].
struct RcSceneBuckets {
    std::uint32_t* count;
    std::uint32_t* start;
    std::uint32_t* blockStart;
    std::uint32_t* cursor;
    std::uint32_t scenes;
};

// Inference with one network per scene: weights of scene s at
// weights + s * kRcParams (BF16); the scene of a query is pathScene[path].
// Queries are sorted by scene into `sorted` (maxCount entries), then
// radiance[path] += throughput[path] * max(out[lane], 0) as LaunchRcInfer.
hipError_t LaunchRcInferScenes(const __bf16* weights, const RcQuery* queries, const std::uint32_t* count,
                               std::uint32_t maxCount, const std::uint32_t* pathScene, const RcSceneBuckets& b,
                               RcQuery* sorted, float* radiance, const float* throughput, unsigned maxBlocks,
                               hipStream_t st);
// He-uniform initial weights for `seed` (bounds sqrt(6 / fan_in), output
// layer scaled by 0.5), FP32, kRcParams values.
void InitRcWeights(std::uint64_t seed, float* w);

// Training step workspace: bf16 rows of RcGradPitch(count) samples each:
// the encoded inputs (128 rows), h1..h5 (5 x 64), the deltas d1..d5
// (5 x 64) and dy as hi and lo planes (6 rows).
constexpr int kRcActX = 0;
constexpr int kRcActH = 128;
constexpr int kRcActD = 128 + 5 * 64;
constexpr int kRcActDy = 128 + 10 * 64;
constexpr int kRcActRows = kRcActDy + 6;
// Partial gradient rows written by LaunchRcGradient (the Adam step sums them).
constexpr int kRcGradSplits = 16;
constexpr std::uint32_t RcGradPitch(std::uint32_t count) { return (count + 31U) / 32U * 32U; }
constexpr std::size_t RcGradWorkspaceBytes(std::uint32_t count)
{
    return static_cast<std::size_t>(kRcActRows) * RcGradPitch(count) * sizeof(std::uint16_t);
}

// Forward + backward over `count` samples (workspace of
// RcGradWorkspaceBytes(count)); writes d(mean loss)/dW as kRcGradSplits
// partial rows of kRcParams floats (overwritten, no accumulation);
// loss[0] += summed loss, loss[1] += samples.
hipError_t LaunchRcGradient(const RcNetwork& net, const RcQuery* samples, std::uint32_t count, __bf16* workspace,
                            float* partial, float* loss, hipStream_t st);
// Adam step (bias-corrected) on all parameters from the sum of `blocks`
// partial rows in s.grad; updates the weight average and refreshes both
// BF16 copies.
hipError_t LaunchRcAdamBlocks(const RcState& s, int blocks, int step, float lr, float beta1, float beta2, float eps,
                              hipStream_t st);
// out[i] = sum of the partial rows.
hipError_t LaunchRcReduce(const float* partial, int blocks, float* out, hipStream_t st);
// loss[0] = loss[1] = 0 (a kernel, not a memset).
hipError_t LaunchRcClearLoss(float* loss, hipStream_t st);
// *out = loss[0] / loss[1] (0 without samples).
hipError_t LaunchRcMeanLoss(const float* loss, float* out, hipStream_t st);
// Resets the average to the FP32 master weights and publishes both BF16 copies.
hipError_t LaunchRcPublish(const RcState& s, hipStream_t st);

} // namespace hdcx
