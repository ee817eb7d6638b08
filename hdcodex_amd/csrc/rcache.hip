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

namespace hdcx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kInferBlock = 256;
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
                    sincospif(ldexpf(pos ? p[a] : d[a - 3], k), &sn, &cs);
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
            x[s][e] = static_cast<__bf16>(v0);
            x[s][e + 1] = static_cast<__bf16>(v1);
        }
    }
}

__device__ __forceinline__ void encode(const RcQuery& q, int h, bf16x8 x[8])
{
    if (h == 0) encodeHalf<0>(q, x);
    else encodeHalf<1>(q, x);
}

__device__ __forceinline__ void reluPack(const f32x16& a, bf16x8& lo, bf16x8& hi)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        lo[e] = static_cast<__bf16>(fmaxf(a[e], 0.0f));
        hi[e] = static_cast<__bf16>(fmaxf(a[e + 8], 0.0f));
    }
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

// Full forward pass of 32 queries; y holds outputs 0..2 in registers 0..2
// of lanes 0..31. Hidden activations optionally stored for backward.
template <bool kStore>
__device__ f32x16 forward(const __bf16* w, const uint4* ldsTiles, const bf16x8 x[8], int lane, __bf16* hs)
{
    const int h = lane >> 5, q = lane & 31;
    f32x16 a0 = {}, a1 = {};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const bf16x8 w0 = ldsTiles ? asBf16x8(ldsTiles[s * 64 + lane]) : weightTile(w, 0, 0, s, lane);
        const bf16x8 w1 = ldsTiles ? asBf16x8(ldsTiles[(8 + s) * 64 + lane]) : weightTile(w, 0, 1, s, lane);
        a0 = mfma(w0, x[s], a0);
        a1 = mfma(w1, x[s], a1);
    }
    bf16x8 hc[4];
    reluPack(a0, hc[0], hc[1]);
    reluPack(a1, hc[2], hc[3]);
    for (int layer = 1; layer <= 4; ++layer) {
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
            const bf16x8 w0 = ldsTiles ? asBf16x8(ldsTiles[(base + s) * 64 + lane]) : weightTile(w, layer, 0, s, lane);
            const bf16x8 w1 = ldsTiles ? asBf16x8(ldsTiles[(base + 4 + s) * 64 + lane]) : weightTile(w, layer, 1, s, lane);
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
    f32x16 y = {};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bf16x8 w5 = ldsTiles ? asBf16x8(ldsTiles[(48 + s) * 64 + lane]) : weightTile(w, 5, 0, s, lane);
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
                                          std::uint32_t end, int lane, float* out, float* radiance, const float* throughput)
{
    const std::uint32_t qi = first + static_cast<std::uint32_t>(lane & 31);
    const bool valid = qi < end;
    RcQuery q{};
    if (valid) q = queries[qi];
    bf16x8 x[8];
    encode(q, lane >> 5, x);
    const f32x16 y = forward<false>(w, tiles, x, lane, nullptr);
    if (valid && lane < 32) {
        if (out) {
            out[qi * 3 + 0] = y[0];
            out[qi * 3 + 1] = y[1];
            out[qi * 3 + 2] = y[2];
        }
        if (radiance) {
            const float v = q.lane == 0 ? y[0] : (q.lane == 1 ? y[1] : y[2]);
            radiance[q.path] += throughput[q.path] * fmaxf(v, 0.0f);
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
                                                            float* out, float* radiance, const float* throughput)
{
    __shared__ uint4 tiles[kWeightTiles * 64];  // 52 KB of MFMA-ready weight tiles
    const std::uint32_t count = countPtr ? min(*countPtr, maxCount) : maxCount;
    // Blocks without a tile of queries leave before staging the weights.
    if (blockIdx.x * (blockDim.x >> 6) >= (count + 31U) / 32U) return;
    stageTiles(tiles, net.w);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const std::uint32_t tileCount = (count + 31U) / 32U;
    for (std::uint32_t tile = wave; tile < tileCount; tile += waves)
        inferTile(net.w, tiles, queries, tile * 32U, count, lane, kOut ? out : nullptr, radiance, throughput);
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

// Forward + backward of tiles of 32 training samples, one wave per block.
// Loss per sample: (y_k - t)^2 / (y_k^2 + 0.01) on the sample's lane k with
// the denominator held constant (relative L2), averaged over the batch.
// Each block accumulates its weight gradient into its own partial row
// partial[blockIdx.x][*] (plain read-modify-write: the row is private).
__global__ __launch_bounds__(64) void rcGradKernel(RcNetwork net, const RcQuery* samples,
                                                  const std::uint32_t* countPtr, std::uint32_t maxCount,
                                                  float* partial, float* loss)
{
    __shared__ __bf16 h0[32 * kP0];
    __shared__ __bf16 hs[5 * 32 * kPH];  // h1..h5
    __shared__ __bf16 ds[5 * 32 * kPH];  // delta1..delta5
    __shared__ float dy[32 * 3];
    const int lane = threadIdx.x, h = lane >> 5, q = lane & 31;
    const std::uint32_t count = countPtr ? min(*countPtr, maxCount) : maxCount;
    const float invN = count ? 1.0f / static_cast<float>(count) : 0.0f;
    float* g = partial + static_cast<std::size_t>(blockIdx.x) * kRcParams;
    const std::uint32_t tileCount = (count + 31U) / 32U;
    float lossSum = 0.0f, lossCount = 0.0f;
    for (std::uint32_t tile = blockIdx.x; tile < tileCount; tile += gridDim.x) {
        const std::uint32_t qi = tile * 32U + static_cast<std::uint32_t>(q);
        const bool valid = qi < count;
        RcQuery smp{};
        if (valid) smp = samples[qi];
        bf16x8 x[8];
        encode(smp, h, x);
#pragma unroll
        for (int s = 0; s < 8; ++s) *reinterpret_cast<bf16x8*>(h0 + q * kP0 + 16 * s + 8 * h) = x[s];
        const f32x16 y = forward<true>(net.w, nullptr, x, lane, hs);
        if (lane < 32) {
            const std::uint32_t k = smp.lane < 3U ? smp.lane : 0U;
            const float yk = k == 0 ? y[0] : (k == 1 ? y[1] : y[2]);
            const float den = yk * yk + 0.01f;
            const float diff = yk - smp.target;
            const float gk = valid ? 2.0f * diff / den * invN : 0.0f;
            dy[q * 3 + 0] = k == 0 ? gk : 0.0f;
            dy[q * 3 + 1] = k == 1 ? gk : 0.0f;
            dy[q * 3 + 2] = k == 2 ? gk : 0.0f;
            if (valid) {
                lossSum += diff * diff / den;
                lossCount += 1.0f;
            }
        }
        __syncthreads();
        // delta5 = (W5^T dy) * relu'(h5): 3 x 64 product on the vector units.
        {
            const __bf16* w5 = net.w + kRcW5;
            const float d0 = dy[q * 3 + 0], d1 = dy[q * 3 + 1], d2 = dy[q * 3 + 2];
            for (int i = 0; i < 32; ++i) {
                const int k = 32 * h + i;
                float v = static_cast<float>(w5[k]) * d0 + static_cast<float>(w5[64 + k]) * d1 +
                          static_cast<float>(w5[128 + k]) * d2;
                if (!(static_cast<float>(hs[4 * 32 * kPH + q * kPH + k]) > 0.0f)) v = 0.0f;
                ds[4 * 32 * kPH + q * kPH + k] = static_cast<__bf16>(v);
            }
        }
        __syncthreads();
        // delta_L = (W_L^T delta_{L+1}) * relu'(h_L), L = 4..1 (W_L: h_L -> h_{L+1}).
        for (int L = 4; L >= 1; --L) {
            const __bf16* wt = net.wt + 4096 * (L - 1);
            const __bf16* dnext = ds + L * 32 * kPH;
            f32x16 a0 = {}, a1 = {};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(dnext + q * kPH + 16 * s + 8 * h);
                const bf16x8 w0 = asBf16x8(*reinterpret_cast<const uint4*>(wt + swap23(q) * 64 + 16 * s + 8 * h));
                const bf16x8 w1 = asBf16x8(*reinterpret_cast<const uint4*>(wt + (32 + swap23(q)) * 64 + 16 * s + 8 * h));
                a0 = mfma(w0, b, a0);
                a1 = mfma(w1, b, a1);
            }
            const __bf16* hl = hs + (L - 1) * 32 * kPH + q * kPH;
            __bf16* dl = ds + (L - 1) * 32 * kPH + q * kPH;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f32x16& a = t == 0 ? a0 : a1;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int k0 = 32 * t + 16 * half + 8 * h;
                    const bf16x8 hv = *reinterpret_cast<const bf16x8*>(hl + k0);
                    bf16x8 out;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float v = a[8 * half + e];
                        out[e] = static_cast<__bf16>(static_cast<float>(hv[e]) > 0.0f ? v : 0.0f);
                    }
                    *reinterpret_cast<bf16x8*>(dl + k0) = out;
                }
            }
            __syncthreads();
        }
        // Weight gradients dW[o][k] = sum_q delta[q][o] in[q][k] with the 32
        // samples as the MFMA reduction dimension (2 k-steps of 16).
        for (int layer = 0; layer <= 4; ++layer) {
            const __bf16* dsrc = ds + layer * 32 * kPH;                    // delta_{layer+1}
            const __bf16* isrc = layer == 0 ? h0 : hs + (layer - 1) * 32 * kPH;
            const int ipitch = layer == 0 ? kP0 : kPH;
            const int kTiles = layer == 0 ? 4 : 2;
            float* gl = g + (layer == 0 ? kRcW0 : kRcW1 + 4096 * (layer - 1));
            const int width = layer == 0 ? kRcIn : kRcHidden;
            for (int to = 0; to < 2; ++to) {
                bf16x8 a[2];
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int e = 0; e < 8; ++e) a[s][e] = dsrc[(16 * s + 8 * h + e) * kPH + 32 * to + q];
                for (int tk = 0; tk < kTiles; ++tk) {
                    f32x16 acc = {};
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        bf16x8 b;
#pragma unroll
                        for (int e = 0; e < 8; ++e) b[e] = isrc[(16 * s + 8 * h + e) * ipitch + 32 * tk + q];
                        acc = mfma(a[s], b, acc);
                    }
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int o = 32 * to + 8 * (j >> 2) + 4 * h + (j & 3);
                        gl[o * width + 32 * tk + q] += acc[j];
                    }
                }
            }
        }
        // dW5[o][k] = sum_q dy[q][o] h5[q][k] on the vector units.
        {
            const int k = lane;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
            for (int i = 0; i < 32; ++i) {
                const float hv = static_cast<float>(hs[4 * 32 * kPH + i * kPH + k]);
                s0 += dy[i * 3 + 0] * hv;
                s1 += dy[i * 3 + 1] * hv;
                s2 += dy[i * 3 + 2] * hv;
            }
            g[kRcW5 + k] += s0;
            g[kRcW5 + 64 + k] += s1;
            g[kRcW5 + 128 + k] += s2;
        }
        __syncthreads();
    }
    // Wave reduction of the loss statistics, one atomic pair per block.
    for (int off = 32; off > 0; off >>= 1) {
        lossSum += __shfl_xor(lossSum, off);
        lossCount += __shfl_xor(lossCount, off);
    }
    if (lane == 0 && loss && lossCount > 0.0f) {
        atomicAdd(&loss[0], lossSum);
        atomicAdd(&loss[1], lossCount);
    }
}

// Sums the per-block partial gradients of one parameter and clears them.
__device__ __forceinline__ float reducePartials(float* partial, int blocks, int i)
{
    float gsum = 0.0f;
    for (int b = 0; b < blocks; ++b) {
        float* p = partial + static_cast<std::size_t>(b) * kRcParams + i;
        gsum += *p;
        *p = 0.0f;
    }
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

__global__ void rcReduceKernel(float* partial, int blocks, float* out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kRcParams; i += gridDim.x * blockDim.x)
        out[i] = reducePartials(partial, blocks, i);
}

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

hipError_t LaunchRcInfer(const RcNetwork& net, const RcQuery* queries, const std::uint32_t* count,
                         std::uint32_t maxCount, float* out, float* radiance, const float* throughput,
                         unsigned maxBlocks, hipStream_t st)
{
    const unsigned tiles = (maxCount + 31U) / 32U;
    unsigned blocks = (tiles + 3U) / 4U;
    blocks = blocks == 0 ? 1U : (blocks < maxBlocks ? blocks : maxBlocks);
    if (out)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(rcInferKernel<true>), dim3(blocks), dim3(kInferBlock), 0, st, net, queries,
                           count, maxCount, out, radiance, throughput);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(rcInferKernel<false>), dim3(blocks), dim3(kInferBlock), 0, st, net, queries,
                           count, maxCount, out, radiance, throughput);
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

hipError_t LaunchRcGradient(const RcNetwork& net, const RcQuery* samples, const std::uint32_t* count,
                            std::uint32_t maxCount, float* grad, float* loss, unsigned maxBlocks, hipStream_t st)
{
    const unsigned tiles = (maxCount + 31U) / 32U;
    const unsigned blocks = tiles == 0 ? 1U : (tiles < maxBlocks ? tiles : maxBlocks);
    hipLaunchKernelGGL(rcGradKernel, dim3(blocks), dim3(64), 0, st, net, samples, count, maxCount, grad, loss);
    return hipGetLastError();
}

hipError_t LaunchRcAdamBlocks(const RcState& s, int blocks, int step, float lr, float beta1, float beta2, float eps,
                              hipStream_t st)
{
    hipLaunchKernelGGL(rcAdamKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, s, blocks, step, lr, beta1,
                       beta2, eps);
    return hipGetLastError();
}

hipError_t LaunchRcReduce(float* partial, int blocks, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(rcReduceKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, partial, blocks, out);
    return hipGetLastError();
}

hipError_t LaunchRcPublish(const RcState& s, hipStream_t st)
{
    hipLaunchKernelGGL(rcPublishKernel, dim3((kRcParams + 255) / 256), dim3(256), 0, st, s);
    return hipGetLastError();
}

} // namespace hdcx
