// Emissive-triangle distribution and dome importance tables on the device
// (see tables.h). FP64 sums in a fixed order: a thread adds its contiguous
// run of kPerThread items in order, a block combines its 256 thread totals
// with a shared-memory scan, and blocks are combined in index order by one
// thread.
#include "tables.h"

#include <algorithm>
#include <cmath>

namespace hdcx {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kPerThread = 16;
constexpr unsigned kPerBlock = kThreads * kPerThread;  // items per block of a scan

// Inclusive scan of one double per thread over the block (Hillis-Steele in
// shared memory); returns this thread's inclusive value; *blockTotal gets
// the block's sum.
__device__ double blockInclusiveScan(double v, double* blockTotal)
{
    __shared__ double buf[kThreads];
    const unsigned t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (unsigned off = 1; off < kThreads; off <<= 1) {
        const double add = t >= off ? buf[t - off] : 0.0;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const double r = buf[t];
    if (blockTotal) *blockTotal = buf[kThreads - 1];
    __syncthreads();
    return r;
}

// ---- Emissive triangles (:2944-2993) ----

__global__ __launch_bounds__(kThreads) void emissiveWeightKernel(const float4* positions, const std::uint32_t* indices,
                                                                 const std::uint32_t* triMaterial, std::uint32_t drawn,
                                                                 const float* power, std::uint32_t materials,
                                                                 double* weight, float* area, double* blockSum)
{
    const std::uint32_t first = blockIdx.x * kPerBlock + threadIdx.x * kPerThread;
    double sum = 0.0;
    for (unsigned k = 0; k < kPerThread; ++k) {
        const std::uint32_t p = first + k;
        if (p >= drawn) break;
        const float4 a = positions[indices[3U * p]], b = positions[indices[3U * p + 1U]], c = positions[indices[3U * p + 2U]];
        const double e1x = static_cast<double>(b.x) - a.x, e1y = static_cast<double>(b.y) - a.y,
                     e1z = static_cast<double>(b.z) - a.z;
        const double e2x = static_cast<double>(c.x) - a.x, e2y = static_cast<double>(c.y) - a.y,
                     e2z = static_cast<double>(c.z) - a.z;
        const double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        float ar = static_cast<float>(0.5 * sqrt(cx * cx + cy * cy + cz * cz));
        ar = isfinite(ar) ? ar : 0.0f;
        const std::uint32_t m = triMaterial[p];
        const float pw = m < materials ? power[m] : 0.0f;
        double w = static_cast<double>(ar) * static_cast<double>(pw);
        if (!(w > 0.0 && isfinite(w))) w = 0.0;
        weight[p] = w;
        area[p] = ar;
        sum += w;
    }
    double total;
    blockInclusiveScan(sum, &total);
    if (threadIdx.x == 0) blockSum[blockIdx.x] = total;
}

// One thread: the total weight, then the offsets of every block in the
// probability scan are computed by emissiveOffsetKernel after the block
// sums of the probabilities.
__global__ void emissiveTotalKernel(const double* blockSum, std::uint32_t blocks, double* total)
{
    double s = 0.0;
    for (std::uint32_t b = 0; b < blocks; ++b) s += blockSum[b];
    *total = s;
}

__device__ __forceinline__ bool usableTotal(double total) { return total > 0.0 && isfinite(total); }

__global__ __launch_bounds__(kThreads) void probabilitySumKernel(const double* weight, std::uint32_t n, const double* total,
                                                                 double* blockSum)
{
    const double tot = *total;
    const std::uint32_t first = blockIdx.x * kPerBlock + threadIdx.x * kPerThread;
    double sum = 0.0;
    if (usableTotal(tot))
        for (unsigned k = 0; k < kPerThread && first + k < n; ++k) sum += weight[first + k] / tot;
    double bt;
    blockInclusiveScan(sum, &bt);
    if (threadIdx.x == 0) blockSum[blockIdx.x] = bt;
}

// Exclusive prefix of the block sums, in block order.
__global__ void blockOffsetKernel(double* blockSum, std::uint32_t blocks)
{
    double s = 0.0;
    for (std::uint32_t b = 0; b < blocks; ++b) {
        const double v = blockSum[b];
        blockSum[b] = s;
        s += v;
    }
}

__global__ __launch_bounds__(kThreads) void emissiveWriteKernel(const double* weight, const float* area, std::uint32_t drawn,
                                                                std::uint32_t all, const double* total,
                                                                const double* blockOffset, float4* out)
{
    const double tot = *total;
    const bool ok = usableTotal(tot);
    const std::uint32_t first = blockIdx.x * kPerBlock + threadIdx.x * kPerThread;
    double local = 0.0;
    if (ok)
        for (unsigned k = 0; k < kPerThread && first + k < drawn; ++k) local += weight[first + k] / tot;
    const double inclusive = blockInclusiveScan(local, nullptr);
    double cdf = (blockIdx.x < (drawn + kPerBlock - 1) / kPerBlock ? blockOffset[blockIdx.x] : 0.0) + inclusive - local;
    for (unsigned k = 0; k < kPerThread; ++k) {
        const std::uint32_t p = first + k;
        if (p >= all) break;
        if (!ok) {
            out[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            continue;
        }
        if (p >= drawn) {
            out[p] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);  // instanced triangles are never sampled as lights
            continue;
        }
        const double probability = weight[p] / tot;
        cdf += probability;
        const float ar = area[p];
        out[p] = make_float4(p + 1U == drawn ? 1.0f : static_cast<float>(cdf),
                             ar > 0.0f ? static_cast<float>(probability) / ar : 0.0f, ar, 0.0f);
    }
}

// ---- Material emission power (:2873-2880 with textureAverageLuminance) ----

__device__ __forceinline__ float srgbToLinear(float c)
{
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// Linear RGB of texel i (clamped to >= 0, non-finite -> 0), as the
// reference's importance and luminance loops read it.
__device__ __forceinline__ void linearTexel(const TexDesc& d, const unsigned char* texels, std::size_t i, float rgb[3])
{
    float v[3];
    if (d.format == HDCX_TEXTURE_RGBA32F) {
        const float4 f = reinterpret_cast<const float4*>(texels + d.offset)[i];
        v[0] = f.x;
        v[1] = f.y;
        v[2] = f.z;
    } else {
        const uchar4 b = reinterpret_cast<const uchar4*>(texels + d.offset)[i];
        v[0] = static_cast<float>(b.x) / 255.0f;
        v[1] = static_cast<float>(b.y) / 255.0f;
        v[2] = static_cast<float>(b.z) / 255.0f;
        if (d.format == HDCX_TEXTURE_RGBA8_SRGB)
            for (int c = 0; c < 3; ++c) v[c] = srgbToLinear(v[c]);
    }
    for (int c = 0; c < 3; ++c) rgb[c] = isfinite(v[c]) ? fmaxf(v[c], 0.0f) : 0.0f;
}

// Mean FP64 luminance of one texture, by the whole block.
__device__ double textureMeanLuminance(const TexDesc& d, const unsigned char* texels)
{
    const std::size_t n = static_cast<std::size_t>(d.width) * d.height;
    double s = 0.0;
    const std::size_t chunk = (n + kThreads - 1) / kThreads;
    const std::size_t b = threadIdx.x * chunk;
    for (std::size_t i = b; i < b + chunk && i < n; ++i) {
        float rgb[3];
        linearTexel(d, texels, i, rgb);
        s += 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2];
    }
    double total;
    blockInclusiveScan(s, &total);
    return n ? total / static_cast<double>(n) : 0.0;
}

__global__ __launch_bounds__(kThreads) void materialPowerKernel(const GpuMaterialRecord* materials, std::uint32_t count,
                                                                const TexDesc* textures, std::uint32_t textureCount,
                                                                const unsigned char* texels, float* power)
{
    const std::uint32_t m = blockIdx.x;
    if (m >= count) return;
    const GpuMaterialRecord& r = materials[m];
    const float weight = fmaxf(r.transmissionOpacityIor.w, 0.0f);
    const std::uint32_t slot = r.textureIndices0.w;
    float luminance;
    if (slot == kMissingTexture) {
        luminance = 0.2126f * fmaxf(r.emissionRoughness.x, 0.0f) + 0.7152f * fmaxf(r.emissionRoughness.y, 0.0f) +
                    0.0722f * fmaxf(r.emissionRoughness.z, 0.0f);
    } else {
        // A plain texture, or a UDIM set (bit 31; first tile in the low byte,
        // tile mask in bits 8..30): the mean of the tiles' averages.
        const bool udim = (slot & 0x80000000U) != 0U;
        const std::uint32_t first = udim ? (slot & 0xffU) : slot;
        const std::uint32_t tiles = udim ? static_cast<std::uint32_t>(__popc((slot >> 8) & 0x7fffffU)) : 1U;
        double total = 0.0;
        for (std::uint32_t k = 0; k < tiles; ++k)
            if (first + k < textureCount) total += static_cast<double>(static_cast<float>(textureMeanLuminance(textures[first + k], texels)));
        luminance = tiles ? static_cast<float>(total / static_cast<double>(tiles)) : 0.0f;
    }
    if (threadIdx.x == 0) power[m] = weight * luminance;
}

// ---- Dome importance (:1998-2066) ----

__device__ __forceinline__ double rowSolidAngle(std::uint32_t y, std::uint32_t w, std::uint32_t h)
{
    const double pi = 3.14159265358979323846;
    const double v0 = static_cast<double>(y) / h, v1 = static_cast<double>(y + 1U) / h;
    return 2.0 * pi / static_cast<double>(w) * fabs(cos(pi * (1.0 - v1)) - cos(pi * (1.0 - v0)));
}

// Per texel: luminance (kept in the pdf plane until the total is known) and
// its weight luminance x row solid angle.
__global__ __launch_bounds__(kThreads) void importanceWeightKernel(TexDesc tex, const unsigned char* texels, float* lum,
                                                                   double* weight)
{
    const std::size_t n = static_cast<std::size_t>(tex.width) * tex.height;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        float rgb[3];
        linearTexel(tex, texels, i, rgb);
        const float l = 0.2126f * rgb[0] + 0.7152f * rgb[1] + 0.0722f * rgb[2];
        lum[i] = l;
        weight[i] = static_cast<double>(l) * rowSolidAngle(static_cast<std::uint32_t>(i / tex.width), tex.width, tex.height);
    }
}

// One block per row: the row's weight, then its conditional CDF.
__global__ __launch_bounds__(kThreads) void importanceRowKernel(const double* weight, std::uint32_t w, double* rowWeight,
                                                                float* conditional)
{
    const std::uint32_t y = blockIdx.x;
    const double* row = weight + static_cast<std::size_t>(y) * w;
    const std::uint32_t chunk = (w + kThreads - 1) / kThreads;
    const std::uint32_t b = threadIdx.x * chunk, e = min(b + chunk, w);
    double s = 0.0;
    for (std::uint32_t x = b; x < e; ++x) s += row[x];
    double rw;
    blockInclusiveScan(s, &rw);
    double local = 0.0;
    for (std::uint32_t x = b; x < e; ++x) local += rw > 0.0 ? row[x] / rw : 1.0 / static_cast<double>(w);
    double cdf = blockInclusiveScan(local, nullptr) - local;
    float* out = conditional + static_cast<std::size_t>(y) * w;
    for (std::uint32_t x = b; x < e; ++x) {
        cdf += rw > 0.0 ? row[x] / rw : 1.0 / static_cast<double>(w);
        out[x] = x + 1U == w ? 1.0f : static_cast<float>(cdf);
    }
    if (threadIdx.x == 0) rowWeight[y] = rw;
}

// One block: the total weight and the marginal CDF.
__global__ __launch_bounds__(kThreads) void importanceMarginalKernel(const double* rowWeight, std::uint32_t h,
                                                                     double* total, float* marginal, std::uint32_t* ok)
{
    const std::uint32_t chunk = (h + kThreads - 1) / kThreads;
    const std::uint32_t b = threadIdx.x * chunk, e = min(b + chunk, h);
    double s = 0.0;
    for (std::uint32_t y = b; y < e; ++y) s += rowWeight[y];
    double tot;
    blockInclusiveScan(s, &tot);
    const bool usable = usableTotal(tot);
    double local = 0.0;
    if (usable)
        for (std::uint32_t y = b; y < e; ++y) local += rowWeight[y] / tot;
    double cdf = blockInclusiveScan(local, nullptr) - local;
    for (std::uint32_t y = b; y < e; ++y) {
        if (usable) cdf += rowWeight[y] / tot;
        marginal[y] = y + 1U == h ? 1.0f : static_cast<float>(cdf);
    }
    if (threadIdx.x == 0) {
        *total = tot;
        *ok = usable ? 1U : 0U;
    }
}

__global__ __launch_bounds__(kThreads) void importancePdfKernel(float* pdf, std::size_t n, const double* total)
{
    const double tot = *total;
    const bool usable = usableTotal(tot);
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
        pdf[i] = usable ? static_cast<float>(static_cast<double>(pdf[i]) / tot) : 0.0f;
}

// Guide entry g of a CDF of n values: the first index whose value is >=
// g / 2^shift (n - 1 when none).
__device__ __forceinline__ std::uint32_t guideEntry(const float* cdf, std::uint32_t n, std::uint32_t g, std::uint32_t shift)
{
    const float threshold = static_cast<float>(static_cast<double>(g) / static_cast<double>(1U << shift));
    std::uint32_t lo = 0, hi = n - 1U;
    while (lo < hi) {
        const std::uint32_t mid = lo + (hi - lo) / 2U;
        if (cdf[mid] >= threshold) hi = mid; else lo = mid + 1U;
    }
    return lo;
}

__global__ __launch_bounds__(kThreads) void importanceGuideKernel(const float* marginal, const float* conditional,
                                                                  std::uint32_t w, std::uint32_t h, std::uint32_t rowShift,
                                                                  std::uint32_t colShift, std::uint32_t* rowGuide,
                                                                  std::uint32_t* colGuide)
{
    const std::uint32_t rowEntries = (1U << rowShift) + 1U, colEntries = (1U << colShift) + 1U;
    const std::size_t n = rowEntries + static_cast<std::size_t>(h) * colEntries;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        if (i < rowEntries) {
            rowGuide[i] = guideEntry(marginal, h, static_cast<std::uint32_t>(i), rowShift);
        } else {
            const std::size_t j = i - rowEntries;
            const std::uint32_t y = static_cast<std::uint32_t>(j / colEntries), g = static_cast<std::uint32_t>(j % colEntries);
            colGuide[j] = guideEntry(conditional + static_cast<std::size_t>(y) * w, w, g, colShift);
        }
    }
}

unsigned GridFor(std::size_t n)
{
    const std::size_t b = (n + kThreads - 1) / kThreads;
    return static_cast<unsigned>(b == 0 ? 1 : (b < 4096 ? b : 4096));
}

} // namespace

std::size_t EmissiveScratchDoubles(std::uint32_t triangles)
{
    const std::size_t blocks = (static_cast<std::size_t>(triangles) + kPerBlock - 1) / kPerBlock + 1;
    // weights, areas (as floats, rounded up to doubles), block sums x 2, total
    return triangles + (triangles + 1) / 2 + 2 * blocks + 2;
}

hipError_t LaunchEmissiveTable(const float4* positions, const std::uint32_t* indices, const std::uint32_t* triMaterial,
                               std::uint32_t drawn, std::uint32_t all, const float* materialPower,
                               std::uint32_t materials, float4* out, double* scratch, hipStream_t st)
{
    const std::uint32_t blocks = (drawn + kPerBlock - 1) / kPerBlock;
    double* weight = scratch;
    float* area = reinterpret_cast<float*>(scratch + drawn);
    double* weightSums = scratch + drawn + (drawn + 1) / 2;
    double* probSums = weightSums + blocks + 1;
    double* total = probSums + blocks + 1;
    if (blocks > 0) {
        hipLaunchKernelGGL(emissiveWeightKernel, dim3(blocks), dim3(kThreads), 0, st, positions, indices, triMaterial, drawn,
                           materialPower, materials, weight, area, weightSums);
    }
    hipLaunchKernelGGL(emissiveTotalKernel, dim3(1), dim3(1), 0, st, weightSums, blocks, total);
    if (blocks > 0) {
        hipLaunchKernelGGL(probabilitySumKernel, dim3(blocks), dim3(kThreads), 0, st, weight, drawn, total, probSums);
        hipLaunchKernelGGL(blockOffsetKernel, dim3(1), dim3(1), 0, st, probSums, blocks);
    }
    const std::uint32_t outBlocks = (all + kPerBlock - 1) / kPerBlock;
    if (outBlocks > 0)
        hipLaunchKernelGGL(emissiveWriteKernel, dim3(outBlocks), dim3(kThreads), 0, st, weight, area, drawn, all, total,
                           probSums, out);
    return hipGetLastError();
}

hipError_t LaunchMaterialPower(const GpuMaterialRecord* materials, std::uint32_t count, const TexDesc* textures,
                               std::uint32_t textureCount, const unsigned char* texels, float* power, hipStream_t st)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(materialPowerKernel, dim3(count), dim3(kThreads), 0, st, materials, count, textures, textureCount,
                       texels, power);
    return hipGetLastError();
}

std::size_t ImportanceScratchDoubles(std::uint32_t width, std::uint32_t height)
{
    return static_cast<std::size_t>(width) * height + height + 1;
}

hipError_t LaunchImportanceTables(const TexDesc& texture, const unsigned char* texels, const ImportanceDesc& d,
                                  float* pool, double* scratch, std::uint32_t* ok, hipStream_t st)
{
    const std::size_t n = static_cast<std::size_t>(d.width) * d.height;
    double* weight = scratch;
    double* rowWeight = scratch + n;
    double* total = rowWeight + d.height;
    float* pdf = pool + d.pdf;
    hipLaunchKernelGGL(importanceWeightKernel, dim3(GridFor(n)), dim3(kThreads), 0, st, texture, texels, pdf, weight);
    hipLaunchKernelGGL(importanceRowKernel, dim3(d.height), dim3(kThreads), 0, st, weight, d.width, rowWeight,
                       pool + d.conditional);
    hipLaunchKernelGGL(importanceMarginalKernel, dim3(1), dim3(kThreads), 0, st, rowWeight, d.height, total,
                       pool + d.marginal, ok);
    hipLaunchKernelGGL(importancePdfKernel, dim3(GridFor(n)), dim3(kThreads), 0, st, pdf, n, total);
    const std::size_t guides = (1U << d.rowShift) + 1U + static_cast<std::size_t>(d.height) * ((1U << d.colShift) + 1U);
    hipLaunchKernelGGL(importanceGuideKernel, dim3(GridFor(guides)), dim3(kThreads), 0, st, pool + d.marginal,
                       pool + d.conditional, d.width, d.height, d.rowShift, d.colShift,
                       reinterpret_cast<std::uint32_t*>(pool + d.rowGuide), reinterpret_cast<std::uint32_t*>(pool + d.colGuide));
    return hipGetLastError();
}

// ---- Shading records (one 128-byte line per triangle) ----
__global__ __launch_bounds__(256) void shadingRecordsKernel(const float4* __restrict__ positions,
                                                            const std::uint32_t* __restrict__ indices,
                                                            const float4* __restrict__ normals,
                                                            const float2* __restrict__ texcoords,
                                                            const std::uint32_t* __restrict__ triMaterial,
                                                            std::uint32_t triangles, float4* __restrict__ out)
{
    for (std::uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < triangles; t += gridDim.x * blockDim.x) {
        const std::size_t c = static_cast<std::size_t>(t) * 3U;
        const float4 p0 = positions[indices[c]], p1 = positions[indices[c + 1]], p2 = positions[indices[c + 2]];
        const float4 n0 = normals[c], n1 = normals[c + 1], n2 = normals[c + 2];
        const float2 u0 = texcoords[c], u1 = texcoords[c + 1], u2 = texcoords[c + 2];
        float4* r = out + static_cast<std::size_t>(t) * 8U;
        r[0] = make_float4(p0.x, p0.y, p0.z, p1.x);
        r[1] = make_float4(p1.y, p1.z, p2.x, p2.y);
        r[2] = make_float4(p2.z, n0.x, n0.y, n0.z);
        r[3] = make_float4(n1.x, n1.y, n1.z, n2.x);
        r[4] = make_float4(n2.y, n2.z, u0.x, u0.y);
        r[5] = make_float4(u1.x, u1.y, u2.x, u2.y);
        r[6] = make_float4(__uint_as_float(triMaterial[t]), 0.0f, 0.0f, 0.0f);
        r[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

hipError_t LaunchShadingRecords(const float4* positions, const std::uint32_t* indices, const float4* normals,
                                const float2* texcoords, const std::uint32_t* triMaterial, std::uint32_t triangles,
                                float4* out, hipStream_t st)
{
    if (triangles == 0) return hipSuccess;
    const unsigned blocks = std::min<unsigned>((triangles + 255U) / 256U, 16384U);
    hipLaunchKernelGGL(shadingRecordsKernel, dim3(blocks), dim3(256), 0, st, positions, indices, normals, texcoords,
                       triMaterial, triangles, out);
    return hipGetLastError();
}

} // namespace hdcx
