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
// radiance[path] += throughput[path] * max(out[lane], 0) when radiance != nullptr.
hipError_t LaunchRcInfer(const RcNetwork& net, const RcQuery* queries, const std::uint32_t* count,
                         std::uint32_t maxCount, float* out, float* radiance, const float* throughput,
                         unsigned maxBlocks, hipStream_t st);
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

// Forward + backward over `count` samples. Block b adds its share of
// d(mean loss)/dW into partial row b (rows are kRcParams floats; at most
// maxBlocks rows are used); loss[0] += summed loss, loss[1] += samples.
hipError_t LaunchRcGradient(const RcNetwork& net, const RcQuery* samples, const std::uint32_t* count,
                            std::uint32_t maxCount, float* partial, float* loss, unsigned maxBlocks, hipStream_t st);
// Adam step (bias-corrected) on all parameters from the sum of `blocks`
// partial rows in s.grad (cleared afterwards); updates the weight average
// and refreshes both BF16 copies.
hipError_t LaunchRcAdamBlocks(const RcState& s, int blocks, int step, float lr, float beta1, float beta2, float eps,
                              hipStream_t st);
// out[i] = sum of the partial rows (cleared afterwards).
hipError_t LaunchRcReduce(float* partial, int blocks, float* out, hipStream_t st);
// Resets the average to the FP32 master weights and publishes both BF16 copies.
hipError_t LaunchRcPublish(const RcState& s, hipStream_t st);

} // namespace hdcx
