// Radiance-cache MLP kernels for gfx950: fused inference and training on
// bf16 MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation, Adam.
//
// Operand convention. Every layer is computed transposed, D^T[o][q] =
// sum_k W[o][k] x[q][k]: the A operand is a 32-row tile of W (rows = output
// units), the B operand the activations of 32 queries (one query per lane
// pair). Lane l of an MFMA result holds query q = l % 32 and, in register
// j, output row 8(j/4) + 4(l/32) + j%4 of the tile. Loading weight row
// 32t + swap23(l % 32) into A-row l % 32 relabels the hidden units so that
// registers j = 0..7 and 8..15 of output tile t are exactly the B operand
// (8 consecutive k) of k-steps 2t and 2t+1 of the next layer: activations
// never leave registers between layers. swap23 exchanges bits 2 and 3.
#include "rcache.h"

#include <cmath>
#include <map>
#include <mutex>

namespace hdcx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Two FP32 values rounded to BF16 as one 32-bit pair (one v_cvt_pk_bf16_f32).
__device__ __forceinline__ std::uint32_t bf16Pair(float a, float b)
{
    return __builtin_bit_cast(std::uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

constexpr int kInferBlock = 512;  // 8 waves share one LDS copy of the weight tiles
constexpr int kWeightTiles = 52;  // 1 KB MFMA-ready A tiles: W0 16, W1..W4 8 each, W5 4
constexpr int kP0 = 136;          // LDS row pitch of the encoded inputs (bf16)
constexpr int kPH = 72;           // LDS row pitch of hidden activations / deltas (bf16)

__device__ __forceinline__ int swap23(int p) { return (p & ~0xC) | ((p & 0x4) << 1) | ((p & 0x8) >> 1); }

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 asBf16x8(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

// Frequency encoding of one query. Logical feature f:
//   [0, 72)    sin/cos(2^k pi p_a), a = f / 24, k = (f % 24) / 2 (k < 12)
//   [72, 108)  sin/cos(2^k pi (d_a + 1) / 2), a = (f-72) / 12, k < 6
//   108 wavelength (lambda0 - 380) / 400, 109 roughness, 110 albedo, 111 one
//   [112, 128) zero
// Lane half h computes features 16 s + 8 h + e (s < 8, e < 8) as the B
// operands of layer 0.
//
// Within a block of 8 features, consecutive bands of one axis follow from
// the previous band by the double-angle formulas (sin 2t = 2 sin t cos t,
// cos 2t = (cos t - sin t)(cos t + sin t)): one sincospif per run of up to
// four bands. Three doublings add at most a few FP32 ulps, far below the
// BF16 rounding of the features.
template <int kH>
__device__ __forceinline__ void encodeHalf(const RcQuery& q, bf16x8 x[8])
{
    const float p[3] = {q.px, q.py, q.pz};
    const float d[3] = {0.5f * (q.dx + 1.0f), 0.5f * (q.dy + 1.0f), 0.5f * (q.dz + 1.0f)};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        int prevAxis = -1, prevBand = -1;  // band carried in (sn, cs), per block
        float sn = 0.0f, cs = 1.0f;
        std::uint32_t pairs[4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const int f = 16 * s + 8 * kH + e;
            float v0, v1;
            if (f < 108) {
                const bool pos = f < 72;
                const int g = pos ? f : f - 72;
                const int a = pos ? g / 24 : 3 + g / 12;  // axes 0-2 position, 3-5 direction
                const int k = pos ? (g % 24) >> 1 : (g % 12) >> 1;
                if (a == prevAxis && k == prevBand + 1) {
                    const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
                    sn = s2;
                    cs = c2;
                } else {
                    // sin and cos of pi y, y = 2^k v (exact), at the hardware
                    // rate: v_sin/v_cos take revolutions, and y / 2 reduced to
                    // [0, 1) by its integer part is exact. Their absolute
                    // error (a few 2^-21) is far below the BF16 rounding of
                    // the features.
                    const float t = 0.5f * ldexpf(pos ? p[a] : d[a - 3], k);
                    const float r = t - floorf(t);
                    sn = __builtin_amdgcn_sinf(r);
                    cs = __builtin_amdgcn_cosf(r);
                }
                prevAxis = a;
                prevBand = k;
                v0 = sn;
                v1 = cs;
            } else if (f == 108) {
                v0 = (q.lambda0 - 380.0f) / 400.0f;
                v1 = q.roughness;
            } else if (f == 110) {
                v0 = q.albedo;
                v1 = 1.0f;
            } else {
                v0 = v1 = 0.0f;
            }
            pairs[e >> 1] = bf16Pair(v0, v1);
        }
        x[s] = __builtin_bit_cast(bf16x8, uint4{pairs[0], pairs[1], pairs[2], pairs[3]});
    }
}

__device__ __forceinline__ void encode(const RcQuery& q, int h, bf16x8 x[8])
{
    if (h == 0) encodeHalf<0>(q, x);
    else encodeHalf<1>(q, x);
}

typedef short i16x8 __attribute__((ext_vector_type(8)));

// BF16 ReLU of a layer's 16 outputs: rounded to BF16 first (v_cvt_pk_bf16_f32,
// two values per instruction), then negative values cleared with a signed
// 16-bit max against 0 on the packed bits (a negative BF16, -0 included, has
// the sign bit set, so it compares below 0). Rounding commutes with
// max(x, 0), so this is max(x, 0) rounded to BF16, in a third of the VALU
// work of an FP32 max (with its NaN canonicalisation) per value.
__device__ __forceinline__ void reluPack(const f32x16& a, bf16x8& lo, bf16x8& hi)
{
    lo = __builtin_bit_cast(bf16x8, uint4{bf16Pair(a[0], a[1]), bf16Pair(a[2], a[3]), bf16Pair(a[4], a[5]),
                                          bf16Pair(a[6], a[7])});
    hi = __builtin_bit_cast(bf16x8, uint4{bf16Pair(a[8], a[9]), bf16Pair(a[10], a[11]), bf16Pair(a[12], a[13]),
                                          bf16Pair(a[14], a[15])});
    const i16x8 zero = {};
    lo = __builtin_bit_cast(bf16x8, __builtin_elementwise_max(__builtin_bit_cast(i16x8, lo), zero));
    hi = __builtin_bit_cast(bf16x8, __builtin_elementwise_max(__builtin_bit_cast(i16x8, hi), zero));
}

__device__ __forceinline__ const __bf16* layerBase(const __bf16* w, int layer)
{
    return w + (layer == 0 ? kRcW0 : (layer == 5 ? kRcW5 : kRcW1 + 4096 * (layer - 1)));
}

// A operand of weight tile (layer, t, s) for this lane, read from global
// row-major weights (rows of W5 beyond the 3 outputs are zero).
__device__ __forceinline__ bf16x8 weightTile(const __bf16* w, int layer, int t, int s, int lane)
{
    const int width = layer == 0 ? kRcIn : kRcHidden;
    const int row = 32 * t + swap23(lane & 31);
    if (layer == 5 && row >= kRcOut) return bf16x8{};
    const __bf16* src = layerBase(w, layer) + row * width + 16 * s + 8 * (lane >> 5);
    return asBf16x8(*reinterpret_cast<const uint4*>(src));
}

// Full forward pass of 32 queries with the block's weight tiles in LDS; y
// holds outputs 0..2 in registers 0..2 of lanes 0..31. Hidden activations
// optionally stored for backward.
// A scheduling barrier between layers keeps each layer's weight-tile reads
// next to its MFMAs: hoisted all together they held 208 VGPRs and spilled.
template <bool kStore>
__device__ __forceinline__ f32x16 forward(const uint4* ldsTiles, const bf16x8 x[8], int lane, __bf16* hs)
{
    const int h = lane >> 5, q = lane & 31;
    f32x16 a0 = {}, a1 = {};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const bf16x8 w0 = asBf16x8(ldsTiles[s * 64 + lane]);
        const bf16x8 w1 = asBf16x8(ldsTiles[(8 + s) * 64 + lane]);
        a0 = mfma(w0, x[s], a0);
        a1 = mfma(w1, x[s], a1);
    }
    bf16x8 hc[4];
    reluPack(a0, hc[0], hc[1]);
    reluPack(a1, hc[2], hc[3]);
    for (int layer = 1; layer <= 4; ++layer) {
        __builtin_amdgcn_sched_barrier(0);
        if (kStore) {
            __bf16* row = hs + (layer - 1) * 32 * kPH + q * kPH;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                *reinterpret_cast<bf16x8*>(row + 32 * t + 8 * h) = hc[2 * t];
                *reinterpret_cast<bf16x8*>(row + 32 * t + 16 + 8 * h) = hc[2 * t + 1];
            }
        }
        const int base = 16 + (layer - 1) * 8;
        a0 = f32x16{};
        a1 = f32x16{};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 w0 = asBf16x8(ldsTiles[(base + s) * 64 + lane]);
            const bf16x8 w1 = asBf16x8(ldsTiles[(base + 4 + s) * 64 + lane]);
            a0 = mfma(w0, hc[s], a0);
            a1 = mfma(w1, hc[s], a1);
        }
        reluPack(a0, hc[0], hc[1]);
        reluPack(a1, hc[2], hc[3]);
    }
    if (kStore) {
        __bf16* row = hs + 4 * 32 * kPH + q * kPH;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            *reinterpret_cast<bf16x8*>(row + 32 * t + 8 * h) = hc[2 * t];
            *reinterpret_cast<bf16x8*>(row + 32 * t + 16 + 8 * h) = hc[2 * t + 1];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    f32x16 y = {};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bf16x8 w5 = asBf16x8(ldsTiles[(48 + s) * 64 + lane]);
        y = mfma(w5, hc[s], y);
    }
    return y;
}

// The block's copy of a network's MFMA-ready weight tiles (52 KB of LDS).
__device__ void stageTiles(uint4* tiles, const __bf16* w)
{
    for (int c = threadIdx.x; c < kWeightTiles * 64; c += blockDim.x) {
        const int b = c >> 6, l = c & 63;
        int layer, t, s;
        if (b < 16) { layer = 0; t = b >> 3; s = b & 7; }
        else if (b < 48) { layer = 1 + ((b - 16) >> 3); t = ((b - 16) & 7) >> 2; s = (b - 16) & 3; }
        else { layer = 5; t = 0; s = b - 48; }
        tiles[c] = __builtin_bit_cast(uint4, weightTile(w, layer, t, s, l));
    }
}

// One wave: queries [first, first + 32) below end.
__device__ __forceinline__ void inferTile(const __bf16* w, const uint4* tiles, const RcQuery* queries, std::uint32_t first,
                                          std::uint32_t end, int lane, float* out, float* radiance, const float* throughput,
                                          float scale = 1.0f)
{
    const std::uint32_t qi = first + static_cast<std::uint32_t>(lane & 31);
    const bool valid = qi < end;
    RcQuery q{};
    if (valid) q = queries[qi];
    bf16x8 x[8];
    encode(q, lane >> 5, x);
    const f32x16 y = forward<false>(tiles, x, lane, nullptr);
    if (valid && lane < 32) {
        if (out) {
            out[qi * 3 + 0] = y[0];
            out[qi * 3 + 1] = y[1];
            out[qi * 3 + 2] = y[2];
        }
        if (radiance) {
            const float v = q.lane == 0 ? y[0] : (q.lane == 1 ? y[1] : y[2]);
            radiance[q.path] += throughput[q.path] * (scale * fmaxf(v, 0.0f));
        }
    }
}

// kOut: also store the three outputs of every query (host-side queries);
// the render path only adds to the paths' radiance and compiles without it
// (fewer registers, two waves per SIMD).
template <bool kOut>
__global__ __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(2))) void rcInferKernel(
    RcNetwork net, const RcQuery* queries,
                                                            const std::uint32_t* countPtr, std::uint32_t maxCount,
                                                            float* out, float* radiance, const float* throughput,
                                                            const float* calibration)
{
    __shared__ uint4 tiles[kWeightTiles * 64];  // 52 KB of MFMA-ready weight tiles
    const std::uint32_t count = countPtr ? min(*countPtr, maxCount) : maxCount;
    const float scale = calibration ? *calibration : 1.0f;
    // Blocks without a tile of queries leave before staging the weights.
    if (blockIdx.x * (blockDim.x >> 6) >= (count + 31U) / 32U) return;
    stageTiles(tiles, net.w);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const std::uint32_t tileCount = (count + 31U) / 32U;
    for (std::uint32_t tile = wave; tile < tileCount; tile += waves)
        inferTile(net.w, tiles, queries, tile * 32U, count, lane, kOut ? out : nullptr, radiance, throughput, scale);
}

// Calibration sums: one FP64 atomic pair per block.
__global__ __launch_bounds__(256) void rcCalibrateSumKernel(const RcQuery* samples, std::uint32_t count,
                                                            const float* outputs, RcCalibration* state)
{
    __shared__ double part[2][4];
    double t = 0.0, y = 0.0;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const RcQuery q = samples[i];
        const std::uint32_t k = q.lane < 3U ? q.lane : 0U;
        t += static_cast<double>(q.target);
        y += static_cast<double>(fmaxf(outputs[3U * i + k], 0.0f));
    }
    for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_xor(t, off);
        y += __shfl_xor(y, off);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wave] = t;
        part[1][wave] = y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = 0.0, ys = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            ts += part[0][w];
            ys += part[1][w];
        }
        atomicAdd(&state->sums[0], ts);
        atomicAdd(&state->sums[1], ys);
    }
}

__global__ void rcCalibrateFinishKernel(RcCalibration* state, float decay)
{
    const double d = static_cast<double>(decay);
    state->acc[0] = d * state->acc[0] + state->sums[0];
    state->acc[1] = d * state->acc[1] + state->sums[1];
    state->sums[0] = 0.0;
    state->sums[1] = 0.0;
    state->scale = state->acc[1] > 0.0 ? static_cast<float>(state->acc[0] / state->acc[1]) : 1.0f;
}

// ---- One network per scene (batch renderer) ----
// Queries are bucketed by the scene of their path (histogram, scan,
// scatter), then each inference block takes kSceneBlockQueries queries of
// one scene and stages that scene's weights.
constexpr std::uint32_t kSceneBlockQueries = 32U * (kInferBlock / 64);

// Scene counts of a block are kept in LDS when the batch has at most
// kLocalScenes scenes: one global atomic per (block, scene) instead of one
// per query (thousands of queries per scene contend otherwise).
constexpr std::uint32_t kLocalScenes = 4096;
constexpr unsigned kBucketBlocks = 512;

__global__ __launch_bounds__(256) void rcSceneHistogramKernel(const RcQuery* queries, const std::uint32_t* countPtr,
                                                             std::uint32_t maxCount, const std::uint32_t* pathScene,
                                                             RcSceneBuckets b)
{
    __shared__ std::uint32_t local[kLocalScenes];
    const std::uint32_t count = min(*countPtr, maxCount);
    const bool inLds = b.scenes <= kLocalScenes;
    if (inLds)
        for (std::uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x) local[s] = 0U;
    __syncthreads();
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const std::uint32_t scene = pathScene[queries[i].path];
        if (inLds) atomicAdd(&local[scene], 1U);
        else atomicAdd(&b.count[scene], 1U);
    }
    __syncthreads();
    if (inLds)
        for (std::uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x)
            if (local[s]) atomicAdd(&b.count[s], local[s]);
}

// One thread: bucket starts, inference-block starts, scatter cursors; the
// histogram is cleared for the next use.
__global__ void rcSceneScanKernel(RcSceneBuckets b)
{
    std::uint32_t q = 0, blocks = 0;
    for (std::uint32_t s = 0; s < b.scenes; ++s) {
        const std::uint32_t n = b.count[s];
        b.start[s] = q;
        b.cursor[s] = q;
        b.blockStart[s] = blocks;
        q += n;
        blocks += (n + kSceneBlockQueries - 1U) / kSceneBlockQueries;
        b.count[s] = 0U;
    }
    b.start[b.scenes] = q;
    b.blockStart[b.scenes] = blocks;
}

// Same grid and visiting order as the histogram: each block first counts
// its queries per scene, reserves that many slots of each scene's bucket
// with one global atomic, then places its queries with LDS atomics.
__global__ __launch_bounds__(256) void rcSceneScatterKernel(const RcQuery* queries, const std::uint32_t* countPtr,
                                                           std::uint32_t maxCount, const std::uint32_t* pathScene,
                                                           RcSceneBuckets b, RcQuery* sorted)
{
    __shared__ std::uint32_t local[kLocalScenes];
    const std::uint32_t count = min(*countPtr, maxCount);
    const std::uint32_t stride = gridDim.x * blockDim.x;
    if (b.scenes > kLocalScenes) {
        for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
            const RcQuery q = queries[i];
            sorted[atomicAdd(&b.cursor[pathScene[q.path]], 1U)] = q;
        }
        return;
    }
    for (std::uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x) local[s] = 0U;
    __syncthreads();
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        atomicAdd(&local[pathScene[queries[i].path]], 1U);
    __syncthreads();
    // local[s] becomes this block's first slot in bucket s.
    for (std::uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x)
        if (local[s]) local[s] = atomicAdd(&b.cursor[s], local[s]);
    __syncthreads();
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const RcQuery q = queries[i];
        sorted[atomicAdd(&local[pathScene[q.path]], 1U)] = q;
    }
}

__global__ __launch_bounds__(kInferBlock) void rcInferScenesKernel(const __bf16* weights, const RcQuery* sorted,
                                                                  RcSceneBuckets b, float* radiance,
                                                                  const float* throughput)
{
    __shared__ uint4 tiles[kWeightTiles * 64];
    const std::uint32_t total = b.blockStart[b.scenes];
    const int lane = threadIdx.x & 63;
    const std::uint32_t waveInBlock = threadIdx.x >> 6;
    std::uint32_t staged = 0xffffffffU;
    for (std::uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        // Scene of this block of queries: the last s with blockStart[s] <= vb.
        std::uint32_t lo = 0, hi = b.scenes - 1U;
        while (lo < hi) {
            const std::uint32_t mid = (lo + hi + 1U) >> 1;
            if (b.blockStart[mid] <= vb) lo = mid; else hi = mid - 1U;
        }
        const __bf16* w = weights + static_cast<std::size_t>(lo) * kRcParams;
        if (lo != staged) {
            __syncthreads();
            stageTiles(tiles, w);
            __syncthreads();
            staged = lo;
        }
        const std::uint32_t first = b.start[lo] + (vb - b.blockStart[lo]) * kSceneBlockQueries + 32U * waveInBlock;
        if (first < b.start[lo + 1U]) inferTile(w, tiles, sorted, first, b.start[lo + 1U], lane, nullptr, radiance, throughput);
    }
}

// ---- Training: forward + backward, then the weight gradients as GEMMs ----
//
// Step 1 (rcForwardBackwardKernel, one wave per tile of 32 samples): the
// encoded inputs, the activations h1..h5 and the deltas d1..d5 (the loss
// gradient with respect to each layer's pre-activation, d5 = (W5^T dy) *
// relu'(h5), d_L = (W_L^T d_{L+1}) * relu'(h_L)) stay in registers in the
// MFMA operand layout (see the operand convention above) and are written
// transposed, one row per unit and one column per sample, into the bf16
// workspace; dy (3 outputs) as two bf16 planes, hi = bf16(dy) and lo =
// bf16(dy - hi). Per sample: loss (y_k - t)^2 / (y_k^2 + 0.01) on the
// sample's lane k, the denominator held constant (relative L2), averaged
// over the batch.
//
// Step 2 (rcWeightGradKernel): dW_l[o][i] = sum_q d_{l+1}[o][q] in_l[i][q]
// (in_0 = the encoded inputs, in_l = h_l) as MFMA tiles with the samples as
// the reduction dimension: 26 output tiles of 32 x 32 (W0: 2 x 4, W1..W4:
// 2 x 2 each, W5: 2 tiles of the 3 x 64 layer from dy's hi and lo planes),
// each split over kRcGradSplits sample ranges. Every (tile, split) wave
// writes its own rows of the partial gradient (no atomics: the sum is
// deterministic); the Adam step adds the kRcGradSplits partial rows.

// The workspace is tiled by 32 samples: tile t holds kRcActRows rows of 32
// consecutive samples (64 bytes), so every row address of a tile is a
// constant offset from the tile's base.
__device__ __forceinline__ __bf16* actTile(__bf16* act, std::uint32_t tile)
{
    return act + static_cast<std::size_t>(tile) * (kRcActRows * 32);
}

// Stores a layer's 64 units held in B-operand layout (hc[2t + half][e] is
// unit 32t + 16 half + 8h + e of sample q) as rows of the tile.
__device__ __forceinline__ void storeUnits(__bf16* tileBase, int firstRow, int q, int h, const bf16x8 (&v)[4])
{
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int half = 0; half < 2; ++half)
#pragma unroll
            for (int e = 0; e < 8; ++e) tileBase[(firstRow + 32 * t + 16 * half + 8 * h + e) * 32 + q] = v[2 * t + half][e];
}

__global__ __launch_bounds__(256) void rcForwardBackwardKernel(RcNetwork net, const RcQuery* samples, std::uint32_t count,
                                                               std::uint32_t pitch, __bf16* act, float* loss)
{
    const int lane = threadIdx.x & 63, h = lane >> 5, q = lane & 31;
    const float invN = count ? 1.0f / static_cast<float>(count) : 0.0f;
    const std::uint32_t tiles = pitch / 32U;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    float lossSum = 0.0f, lossCount = 0.0f;
    for (std::uint32_t tile = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; tile < tiles; tile += waves) {
        const std::uint32_t col = tile * 32U + static_cast<std::uint32_t>(q);
        const bool valid = col < count;
        RcQuery smp{};
        if (valid) smp = samples[col];
        __bf16* tb = actTile(act, tile);
        bf16x8 x[8];
        encode(smp, h, x);
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) tb[(kRcActX + 16 * s + 8 * h + e) * 32 + q] = x[s][e];
        // Forward. Each hidden layer is stored, and kept as a 32-bit mask of
        // its positive units (bit 8 (2t + half) + e: the relu' of the backward pass).
        bf16x8 hc[4];
        std::uint32_t mask[5];
        {
            f32x16 a0 = {}, a1 = {};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                a0 = mfma(weightTile(net.w, 0, 0, s, lane), x[s], a0);
                a1 = mfma(weightTile(net.w, 0, 1, s, lane), x[s], a1);
            }
            reluPack(a0, hc[0], hc[1]);
            reluPack(a1, hc[2], hc[3]);
        }
#pragma unroll
        for (int layer = 0; layer < 5; ++layer) {
            if (layer > 0) {
                f32x16 a0 = {}, a1 = {};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    a0 = mfma(weightTile(net.w, layer, 0, s, lane), hc[s], a0);
                    a1 = mfma(weightTile(net.w, layer, 1, s, lane), hc[s], a1);
                }
                reluPack(a0, hc[0], hc[1]);
                reluPack(a1, hc[2], hc[3]);
            }
            std::uint32_t m = 0;
#pragma unroll
            for (int b = 0; b < 32; ++b) m |= (static_cast<float>(hc[b >> 3][b & 7]) > 0.0f ? 1U : 0U) << b;
            mask[layer] = m;
            storeUnits(tb, kRcActH + 64 * layer, q, h, hc);
        }
        f32x16 y = {};
#pragma unroll
        for (int s = 0; s < 4; ++s) y = mfma(weightTile(net.w, 5, 0, s, lane), hc[s], y);
        // dy on lanes 0..31 (outputs 0..2 are registers 0..2 there), shared
        // with lanes 32..63 of the same sample.
        float dy0 = 0.0f, dy1 = 0.0f, dy2 = 0.0f;
        if (h == 0) {
            const std::uint32_t k = smp.lane < 3U ? smp.lane : 0U;
            const float yk = k == 0 ? y[0] : (k == 1 ? y[1] : y[2]);
            const float den = yk * yk + 0.01f;
            const float diff = yk - smp.target;
            const float gk = valid ? 2.0f * diff / den * invN : 0.0f;
            dy0 = k == 0 ? gk : 0.0f;
            dy1 = k == 1 ? gk : 0.0f;
            dy2 = k == 2 ? gk : 0.0f;
            if (valid) {
                lossSum += diff * diff / den;
                lossCount += 1.0f;
            }
            const float dv[3] = {dy0, dy1, dy2};
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const __bf16 hi = static_cast<__bf16>(dv[o]);
                tb[(kRcActDy + o) * 32 + q] = hi;
                tb[(kRcActDy + 3 + o) * 32 + q] = static_cast<__bf16>(dv[o] - static_cast<float>(hi));
            }
        }
        dy0 = __shfl(dy0, q);
        dy1 = __shfl(dy1, q);
        dy2 = __shfl(dy2, q);
        // d5 = (W5^T dy) * relu'(h5) in FP32 on the vector units.
        bf16x8 d[4];
        {
            const __bf16* w5 = net.w + kRcW5;
#pragma unroll
            for (int b = 0; b < 32; ++b) {
                const int u = 32 * (b >> 4) + 16 * ((b >> 3) & 1) + 8 * h + (b & 7);
                const float v = static_cast<float>(w5[u]) * dy0 + static_cast<float>(w5[64 + u]) * dy1 +
                                static_cast<float>(w5[128 + u]) * dy2;
                d[b >> 3][b & 7] = static_cast<__bf16>((mask[4] >> b) & 1U ? v : 0.0f);
            }
        }
        storeUnits(tb, kRcActD + 64 * 4, q, h, d);
        // d_L = (W_L^T d_{L+1}) * relu'(h_L), L = 4..1 (W_L maps h_L to h_{L+1}).
#pragma unroll
        for (int L = 4; L >= 1; --L) {
            const __bf16* wt = net.wt + 4096 * (L - 1);
            f32x16 a0 = {}, a1 = {};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bf16x8 w0 = asBf16x8(*reinterpret_cast<const uint4*>(wt + swap23(q) * 64 + 16 * s + 8 * h));
                const bf16x8 w1 = asBf16x8(*reinterpret_cast<const uint4*>(wt + (32 + swap23(q)) * 64 + 16 * s + 8 * h));
                a0 = mfma(w0, d[s], a0);
                a1 = mfma(w1, d[s], a1);
            }
#pragma unroll
            for (int b = 0; b < 32; ++b) {
                const float v = (b >> 4) == 0 ? a0[b & 15] : a1[b & 15];
                d[b >> 3][b & 7] = static_cast<__bf16>((mask[L - 1] >> b) & 1U ? v : 0.0f);
            }
            storeUnits(tb, kRcActD + 64 * (L - 1), q, h, d);
        }
    }
    // Wave reduction of the loss statistics, one atomic pair per wave.
    for (int off = 32; off > 0; off >>= 1) {
        lossSum += __shfl_xor(lossSum, off);
        lossCount += __shfl_xor(lossCount, off);
    }
    if (lane == 0 && loss && lossCount > 0.0f) {
        atomicAdd(&loss[0], lossSum);
        atomicAdd(&loss[1], lossCount);
    }
}

// Output tiles of the weight gradient: 24 tiles of W0..W4, then 2 of W5.
constexpr int kGradTiles = 26;

__global__ __launch_bounds__(256) void rcWeightGradKernel(const __bf16* act, std::uint32_t pitch, float* partial)
{
    __shared__ float w5acc[4][6][32];
    const int lane = threadIdx.x & 63, h = lane >> 5, r = lane & 31, waveInBlock = threadIdx.x >> 6;
    const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int tile = wave % kGradTiles, split = wave / kGradTiles;
    if (split >= kRcGradSplits) return;  // whole waves: no barrier below is shared across waves
    int aRow, bRow, width, out, oBase, iBase;
    bool w5 = false;
    if (tile < 8) {  // W0: 64 x 128
        oBase = 32 * (tile >> 2);
        iBase = 32 * (tile & 3);
        aRow = kRcActD + oBase;
        bRow = kRcActX + iBase;
        width = kRcIn;
        out = kRcW0;
    } else if (tile < 24) {  // W1..W4: 64 x 64
        const int l = 1 + (tile - 8) / 4, t = (tile - 8) & 3;
        oBase = 32 * (t >> 1);
        iBase = 32 * (t & 1);
        aRow = kRcActD + 64 * l + oBase;
        bRow = kRcActH + 64 * (l - 1) + iBase;
        width = kRcHidden;
        out = kRcW1 + 4096 * (l - 1);
    } else {  // W5: 3 x 64 from the hi and lo planes of dy
        w5 = true;
        oBase = 0;
        iBase = 32 * (tile - 24);
        aRow = kRcActDy;
        bRow = kRcActH + 64 * 4 + iBase;
        width = kRcHidden;
        out = kRcW5;
    }
    const std::uint32_t steps = pitch / 16U;
    const std::uint32_t s0 = steps * static_cast<std::uint32_t>(split) / kRcGradSplits;
    const std::uint32_t s1 = steps * static_cast<std::uint32_t>(split + 1) / kRcGradSplits;
    // Step s covers samples 16 s .. 16 s + 15: half (s & 1) of tile s / 2.
    const bool aValid = !w5 || r < 6;
    const std::size_t aOff = static_cast<std::size_t>(aRow + (aValid ? r : 0)) * 32 + 8 * h;
    const std::size_t bOff = static_cast<std::size_t>(bRow + r) * 32 + 8 * h;
    f32x16 acc = {};
    for (std::uint32_t s = s0; s < s1; ++s) {
        const __bf16* tb = act + static_cast<std::size_t>(s >> 1) * (kRcActRows * 32) + 16 * (s & 1U);
        const bf16x8 a = aValid ? asBf16x8(*reinterpret_cast<const uint4*>(tb + aOff)) : bf16x8{};
        const bf16x8 b = asBf16x8(*reinterpret_cast<const uint4*>(tb + bOff));
        acc = mfma(a, b, acc);
    }
    float* g = partial + static_cast<std::size_t>(split) * kRcParams + out;
    if (!w5) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int o = oBase + 8 * (j >> 2) + 4 * h + (j & 3);
            g[o * width + iBase + r] = acc[j];
        }
        return;
    }
    // W5: rows 0..2 (hi) and 3..5 (lo) of the tile, combined through LDS
    // (row 8 (j/4) + 4h + j%4 of register j).
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int o = 8 * (j >> 2) + 4 * h + (j & 3);
        if (o < 6) w5acc[waveInBlock][o][r] = acc[j];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (h == 0)
#pragma unroll
        for (int o = 0; o < 3; ++o)
            g[o * width + iBase + r] = w5acc[waveInBlock][o][r] + w5acc[waveInBlock][3 + o][r];
}

// Sums the partial gradient rows of one parameter (in row order).
__device__ __forceinline__ float reducePartials(const float* partial, int blocks, int i)
{
    float gsum = 0.0f;
    for (int b = 0; b < blocks; ++b) gsum += partial[static_cast<std::size_t>(b) * kRcParams + i];
    return gsum;
}

__device__ __forceinline__ void publish(__bf16* w, __bf16* wt, int i, float value)
{
    w[i] = static_cast<__bf16>(value);
    if (i >= kRcW1 && i < kRcW5) {
        const int r = i - kRcW1, layer = r / 4096, o = (r % 4096) / 64, k = r % 64;
        wt[layer * 4096 + k * 64 + o] = static_cast<__bf16>(value);
    }
}

__global__ void rcAdamKernel(RcState s, int blocks, int step, float lr, float beta1, float beta2, float eps)
{
    const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
    const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kRcParams; i += gridDim.x * blockDim.x) {
        const float gi = reducePartials(s.grad, blocks, i);
        const float m = beta1 * s.m[i] + (1.0f - beta1) * gi;
        const float v = beta2 * s.v[i] + (1.0f - beta2) * gi * gi;
        s.m[i] = m;
        s.v[i] = v;
        const float w = s.master[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
        s.master[i] = w;
        publish(s.w, s.wt, i, w);
        // Moving average started from the initial weights (an average started
        // from zero and bias-corrected was measured 20-30% worse on the
        // heightfield cache comparison).
        const float e = s.emaDecay * s.ema[i] + (1.0f - s.emaDecay) * w;
        s.ema[i] = e;
        publish(s.we, s.wte, i, e);
    }
}

__global__ void rcReduceKernel(const float* partial, int blocks, float* out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kRcParams; i += gridDim.x * blockDim.x)
        out[i] = reducePartials(partial, blocks, i);
}

__global__ void rcClearLossKernel(float* loss)
{
    loss[0] = 0.0f;
    loss[1] = 0.0f;
}

__global__ void rcMeanLossKernel(const float* loss, float* out) { *out = loss[1] > 0.0f ? loss[0] / loss[1] : 0.0f; }

__global__ void rcPublishKernel(RcState s)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kRcParams; i += gridDim.x * blockDim.x) {
        const float w = s.master[i];
        s.ema[i] = w;
        publish(s.w, s.wt, i, w);
        publish(s.we, s.wte, i, w);
    }
}

} // namespace

void InitRcWeights(std::uint64_t seed, float* w)
{
    std::uint64_t state = seed ^ 0x5deece66dULL;
    auto next = [&]() {
        std::uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return static_cast<double>((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < kRcParams; ++i) {
        const int fanIn = i < kRcW1 ? kRcIn : kRcHidden;
        const double bound = std::sqrt(6.0 / fanIn) * (i >= kRcW5 ? 0.5 : 1.0);
        w[i] = static_cast<float>((2.0 * next() - 1.0) * bound);
    }
}

// Blocks of rcInferKernel<kOut> the current device keeps resident (its
// occupancy per CU times the CU count; 0 when the query fails), per device.
template <bool kOut>
unsigned ResidentInferBlocks()
{
    static std::mutex mutex;
    static std::map<int, unsigned> cache;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return 0U;
    const std::lock_guard<std::mutex> lock(mutex);
    if (const auto found = cache.find(device); found != cache.end()) return found->second;
    int perCu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(rcInferKernel<kOut>),
                                                     kInferBlock, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0U;
    const unsigned resident = static_cast<unsigned>(perCu > 0 ? perCu : 1) * static_cast<unsigned>(cus > 0 ? cus : 1);
    cache.emplace(device, resident);
    return resident;
}

hipError_t LaunchRcInfer(const RcNetwork& net, const RcQuery* queries, const std::uint32_t* count,
                         std::uint32_t maxCount, float* out, float* radiance, const float* throughput,
                         unsigned maxBlocks, hipStream_t st, const float* calibration)
{
    // One wave per 32 queries, grid-stride over the tiles, with at most the
    // blocks the device keeps resident: every wave has about the same work,
    // so a partial second round of blocks would idle most of the GPU.
    const unsigned tiles = (maxCount + 31U) / 32U;
    constexpr unsigned kWaves = kInferBlock / 64;
    unsigned blocks = (tiles + kWaves - 1U) / kWaves;
    blocks = blocks == 0 ? 1U : (blocks < maxBlocks ? blocks : maxBlocks);
    const unsigned resident = out ? ResidentInferBlocks<true>() : ResidentInferBlocks<false>();
    if (resident > 0U && blocks > resident) blocks = resident;
    if (out)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(rcInferKernel<true>), dim3(blocks), dim3(kInferBlock), 0, st, net, queries,
                           count, maxCount, out, radiance, throughput, calibration);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(rcInferKernel<false>), dim3(blocks), dim3(kInferBlock), 0, st, net, queries,
                           count, maxCount, out, radiance, throughput, calibration);
    return hipGetLastError();
}

hipError_t LaunchRcCalibrate(const RcQuery* samples, std::uint32_t count, const float* outputs, RcCalibration* state,
                             float decay, hipStream_t st)
{
    const unsigned blocks = count == 0 ? 1U : ((count + 255U) / 256U < 256U ? (count + 255U) / 256U : 256U);
    hipLaunchKernelGGL(rcCalibrateSumKernel, dim3(blocks), dim3(256), 0, st, samples, count, outputs, state);
    hipLaunchKernelGGL(rcCalibrateFinishKernel, dim3(1), dim3(1), 0, st, state, decay);
    return hipGetLastError();
}

hipError_t LaunchRcInferScenes(const __bf16* weights, const RcQuery* queries, const std::uint32_t* count,
                               std::uint32_t maxCount, const std::uint32_t* pathScene, const RcSceneBuckets& b,
                               RcQuery* sorted, float* radiance, const float* throughput, unsigned maxBlocks,
                               hipStream_t st)
{
    const unsigned spread = (maxCount + 255U) / 256U;
    const unsigned cap = maxBlocks < kBucketBlocks ? maxBlocks : kBucketBlocks;
    const unsigned grid = spread == 0 ? 1U : (spread < cap ? spread : cap);
    hipLaunchKernelGGL(rcSceneHistogramKernel, dim3(grid), dim3(256), 0, st, queries, count, maxCount, pathScene, b);
    hipLaunchKernelGGL(rcSceneScanKernel, dim3(1), dim3(1), 0, st, b);
    hipLaunchKernelGGL(rcSceneScatterKernel, dim3(grid), dim3(256), 0, st, queries, count, maxCount, pathScene, b, sorted);
    const unsigned blocks = (maxCount + kSceneBlockQueries - 1U) / kSceneBlockQueries + b.scenes;
    const unsigned inferGrid = blocks < maxBlocks ? blocks : maxBlocks;
    hipLaunchKernelGGL(rcInferScenesKernel, dim3(inferGrid == 0 ? 1U : inferGrid), dim3(kInferBlock), 0, st, weights, sorted,
                       b, radiance, throughput);
    return hipGetLastError();
}

hipError_t LaunchRcGradient(const RcNetwork& net, const RcQuery* samples, std::uint32_t count, __bf16* workspace,
                            float* partial, float* loss, hipStream_t st)
{
    const std::uint32_t pitch = RcGradPitch(count);
    const unsigned tiles = pitch / 32U;
    const unsigned blocks = tiles == 0 ? 1U : (tiles + 3U) / 4U;
    hipLaunchKernelGGL(rcForwardBackwardKernel, dim3(blocks), dim3(256), 0, st, net, samples, count, pitch, workspace, loss);
    const unsigned waves = kGradTiles * kRcGradSplits;
    hipLaunchKernelGGL(rcWeightGradKernel, dim3((waves + 3U) / 4U), dim3(256), 0, st, workspace, pitch, partial);
    return hipGetLastError();
}

hipError_t LaunchRcAdamBlocks(const RcState& s, int blocks, int step, float lr, float beta1, float beta2, float eps,
                              hipStream_t st)
{
    hipLaunchKernelGGL(rcAdamKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, s, blocks, step, lr, beta1,
                       beta2, eps);
    return hipGetLastError();
}

hipError_t LaunchRcReduce(const float* partial, int blocks, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(rcReduceKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, partial, blocks, out);
    return hipGetLastError();
}

hipError_t LaunchRcClearLoss(float* loss, hipStream_t st)
{
    hipLaunchKernelGGL(rcClearLossKernel, dim3(1), dim3(1), 0, st, loss);
    return hipGetLastError();
}

hipError_t LaunchRcMeanLoss(const float* loss, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(rcMeanLossKernel, dim3(1), dim3(1), 0, st, loss, out);
    return hipGetLastError();
}

hipError_t LaunchRcPublish(const RcState& s, hipStream_t st)
{
    hipLaunchKernelGGL(rcPublishKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, s);
    return hipGetLastError();
}

} // namespace hdcx
