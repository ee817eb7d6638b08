// ORACLE — CPU restatement of the reference integrator. TEST INFRASTRUCTURE
// ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker; never linked into or called by the
// product library.
//
// It restates the GLSL compute shader embedded in the reference
// (src/gpu/vulkan_path_tracer.cpp:100-1891) function by function; each
// function cites the lines it follows. The arithmetic type R is float (the
// FP32 restatement used as the CPU baseline) or double (the FP64 oracle).
// Random numbers are produced in FP32 exactly as the shader does and then
// promoted; the hero wavelengths are computed in FP32 as well.
//
// What the reference leaves to the Vulkan driver is defined here and
// documented as such (parity against the vendor binary is unpinned):
//   * ray/triangle intersection: watertight test (Woop, Benthin, Wald,
//     JCGT 2013) with FP contraction off; a hit counts when tmin <= t <= tmax;
//     closest hit = smallest (t, primitive index) pair;
//   * any-hit order in non-opaque scenes: candidates in increasing
//     (t, primitive) order; shadow falloff uses the closest occluder;
//   * texture filtering: bilinear with REPEAT and float weights, sRGB texels
//     decoded with the IEC 61966-2-1 formula before filtering.
#include "oracle_math.hpp"
#include "../include/hdcx.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace oracle {
namespace {

constexpr std::uint32_t kMissing = 0xffffffffU;
constexpr double kPiD = 3.14159265359;  // the shader's literal

struct F4 { float x, y, z, w; };
struct U4 { std::uint32_t x, y, z, w; };

// std430 GpuMaterial (vulkan_path_tracer.cpp:42-66).
struct Material {
    F4 baseColorMetalness, emissionRoughness, transmissionOpacityIor,
        transmissionColorThinWalled, subsurfaceWeightScale, subsurfaceColor,
        subsurfaceRadius, specularWeightColor, coatWeightRoughnessIor, coatColor,
        transmissionMedium, transmissionScatter, sheenColorMode, diffuseProperties,
        translucentColorWeight, baseColorTextureScale, baseColorTextureBias;
    U4 textureIndices0, textureIndices1, textureIndices2, textureIndices3,
        textureIndices4, textureIndices5;
};
static_assert(sizeof(Material) == 368, "GpuMaterial");

// std430 GpuLight (vulkan_path_tracer.cpp:71-86).
struct Light {
    F4 colorIntensity, controls, position, axisU, axisV, basisX, basisY, basisZ,
        shaping, focusTint, shadow, shadowColor, geometry;
    U4 textureInfo;
};
static_assert(sizeof(Light) == 224, "GpuLight");

struct Texture {
    int width = 0, height = 0;
    std::vector<float> texels;  // decoded linear RGBA, float
};

struct Hit {
    bool found = false;
    std::uint32_t primitive = kMissing;
    double t = 0.0;  // in R, stored widened
    double u = 0.0, v = 0.0;
};

struct BvhNode {
    float lo[3], hi[3];
    std::uint32_t left;   // inner: left child; leaf: first index
    std::uint32_t count;  // 0 = inner (right = left + 1 .. stored below)
    std::uint32_t right;
};

float SrgbDecode(float c)
{
    return c <= 0.04045F ? c / 12.92F : std::pow((c + 0.055F) / 1.055F, 2.4F);
}

class Scene {
public:
    explicit Scene(const hdcx_scene_view& v)
    {
        vertexCount = v.vertex_count;
        triangleCount = v.triangle_count;
        positions.assign(v.positions, v.positions + static_cast<std::size_t>(vertexCount) * 4);
        indices.assign(v.indices, v.indices + static_cast<std::size_t>(triangleCount) * 3);
        triangleMaterials.assign(v.triangle_materials, v.triangle_materials + triangleCount);
        texcoords.assign(v.texcoords, v.texcoords + static_cast<std::size_t>(triangleCount) * 6);
        normals.assign(v.normals, v.normals + static_cast<std::size_t>(triangleCount) * 12);
        if (v.emissive) emissive.assign(v.emissive, v.emissive + static_cast<std::size_t>(triangleCount) * 4);
        materials.resize(v.material_count);
        std::memcpy(materials.data(), v.materials, sizeof(Material) * v.material_count);
        lightCount = v.light_count;
        lights.resize(std::max<std::uint32_t>(v.light_count, 1));
        std::memcpy(lights.data(), v.lights, sizeof(Light) * lights.size());
        opaque = v.opaque != 0;
        for (std::uint32_t i = 0; i < v.texture_count; ++i) {
            const hdcx_texture_view& tv = v.textures[i];
            Texture t;
            t.width = static_cast<int>(tv.width);
            t.height = static_cast<int>(tv.height);
            const std::size_t n = static_cast<std::size_t>(tv.width) * tv.height * 4;
            t.texels.resize(n);
            if (tv.format == HDCX_TEXTURE_RGBA32F) {
                std::memcpy(t.texels.data(), tv.texels, n * sizeof(float));
            } else {
                const auto* b = static_cast<const std::uint8_t*>(tv.texels);
                for (std::size_t k = 0; k < n; ++k) {
                    float c = static_cast<float>(b[k]) / 255.0F;
                    if (tv.format == HDCX_TEXTURE_RGBA8_SRGB && (k & 3) != 3) c = SrgbDecode(c);
                    t.texels[k] = c;
                }
            }
            textures.push_back(std::move(t));
        }
        if (!v.emissive) BuildEmissive(v);
        BuildDomeImportance();
        BuildBvh();
    }

    std::uint32_t vertexCount = 0, triangleCount = 0, lightCount = 0;
    std::vector<float> positions, texcoords, normals, emissive;
    std::vector<std::uint32_t> indices, triangleMaterials;
    std::vector<Material> materials;
    std::vector<Light> lights;
    std::vector<Texture> textures;
    bool opaque = true;

    std::vector<BvhNode> nodes;
    std::vector<std::uint32_t> order;

    template <class R> V3<R> Pos(std::uint32_t i) const
    {
        return {R(positions[i * 4 + 0]), R(positions[i * 4 + 1]), R(positions[i * 4 + 2])};
    }

private:
    static void LinearRgb(const Texture& t, std::size_t texel, float rgb[3])
    {
        for (int c = 0; c < 3; ++c) {
            const float x = t.texels[texel * 4 + c];
            rgb[c] = std::isfinite(x) ? std::max(x, 0.0F) : 0.0F;
        }
    }
    // TextureAverageLuminance (:1968-1996): FP64 mean of the texels'
    // luminance, returned as float.
    float AverageLuminance(std::uint32_t ti) const
    {
        const Texture& t = textures[ti];
        const std::size_t n = static_cast<std::size_t>(t.width) * t.height;
        double total = 0.0;
        for (std::size_t p = 0; p < n; ++p) {
            float rgb[3];
            LinearRgb(t, p, rgb);
            total += 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2];
        }
        return n ? static_cast<float>(total / static_cast<double>(n)) : 0.0F;
    }
    // materialEmissionPowers (:2873-2880) from a packed record: the
    // emission luminance, or the emission texture's mean luminance (a UDIM
    // set: the mean of its tiles' means), times max(emissionWeight, 0).
    float MaterialPower(const Material& m) const
    {
        const std::uint32_t slot = m.textureIndices0.w;
        float luminance;
        if (slot == kMissing) {
            luminance = 0.2126F * std::max(m.emissionRoughness.x, 0.0F) + 0.7152F * std::max(m.emissionRoughness.y, 0.0F) +
                        0.0722F * std::max(m.emissionRoughness.z, 0.0F);
        } else {
            const bool udim = (slot & 0x80000000U) != 0U;
            const std::uint32_t first = udim ? (slot & 0xffU) : slot;
            const std::uint32_t tiles = udim ? static_cast<std::uint32_t>(__builtin_popcount((slot >> 8) & 0x7fffffU)) : 1U;
            double total = 0.0;
            for (std::uint32_t k = 0; k < tiles; ++k) total += AverageLuminance(first + k);
            luminance = static_cast<float>(total / static_cast<double>(tiles));
        }
        return std::max(m.transmissionOpacityIor.w, 0.0F) * luminance;
    }
    // The emissive-triangle distribution of BuildScene (:2944-2993): weight
    // = area x material power (area from the FP64 cross product of the FP32
    // positions), FP64 sequential sums, cdf of the last sampled triangle 1;
    // triangles past the sampled ones: cdf 1, pdf 0; all zero without a
    // positive weight.
    void BuildEmissive(const hdcx_scene_view& v)
    {
        const std::uint32_t drawn = v.prototype_count ? v.prototypes[0].first_triangle : triangleCount;
        const std::uint32_t n = v.light_triangle_count ? v.light_triangle_count : drawn;
        std::vector<float> power(materials.size());
        for (std::size_t m = 0; m < materials.size(); ++m)
            power[m] = v.material_emission_power ? v.material_emission_power[m] : MaterialPower(materials[m]);
        emissive.assign(static_cast<std::size_t>(triangleCount) * 4, 0.0F);
        std::vector<double> weight(n, 0.0);
        std::vector<float> area(n, 0.0F);
        double total = 0.0;
        for (std::uint32_t p = 0; p < n; ++p) {
            const float* a = &positions[static_cast<std::size_t>(indices[3 * p]) * 4];
            const float* b = &positions[static_cast<std::size_t>(indices[3 * p + 1]) * 4];
            const float* c = &positions[static_cast<std::size_t>(indices[3 * p + 2]) * 4];
            double e1[3], e2[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = static_cast<double>(b[k]) - a[k];
                e2[k] = static_cast<double>(c[k]) - a[k];
            }
            const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2],
                         z = e1[0] * e2[1] - e1[1] * e2[0];
            const float ar = static_cast<float>(0.5 * std::sqrt(x * x + y * y + z * z));
            area[p] = std::isfinite(ar) ? ar : 0.0F;
            const std::uint32_t m = triangleMaterials[p];
            const double w = static_cast<double>(area[p]) * static_cast<double>(m < power.size() ? power[m] : 0.0F);
            if (w > 0.0 && std::isfinite(w)) {
                weight[p] = w;
                total += w;
            }
        }
        if (!(total > 0.0 && std::isfinite(total))) return;
        double cdf = 0.0;
        for (std::uint32_t p = 0; p < n; ++p) {
            const double q = weight[p] / total;
            cdf += q;
            emissive[static_cast<std::size_t>(p) * 4] = p + 1 == n ? 1.0F : static_cast<float>(cdf);
            emissive[static_cast<std::size_t>(p) * 4 + 1] = area[p] > 0.0F ? static_cast<float>(q) / area[p] : 0.0F;
            emissive[static_cast<std::size_t>(p) * 4 + 2] = area[p];
        }
        for (std::uint32_t p = n; p < triangleCount; ++p) emissive[static_cast<std::size_t>(p) * 4] = 1.0F;
    }
    // Lat-long importance tables (:1998-2066) of the texture of every dome
    // whose importance slot names its own texture, appended as a texture
    // (r conditional CDF, g marginal CDF, b pdf per steradian) that the
    // dome's slot then names; without positive weight the dome is sampled
    // uniformly.
    void BuildDomeImportance()
    {
        std::vector<std::pair<std::uint32_t, std::uint32_t>> built;  // source texture -> table texture
        for (std::uint32_t l = 0; l < lightCount; ++l) {
            U4& info = lights[l].textureInfo;
            if (info.z != 0U || info.w == kMissing || info.w != info.x) continue;
            std::uint32_t table = kMissing;
            bool found = false;
            for (const auto& b : built)
                if (b.first == info.x) { table = b.second; found = true; }
            if (!found) {
                table = ImportanceTable(textures[info.x]);
                built.emplace_back(info.x, table);
            }
            info.w = table;
        }
    }
    std::uint32_t ImportanceTable(const Texture& src)
    {
        const std::uint32_t w = static_cast<std::uint32_t>(src.width), h = static_cast<std::uint32_t>(src.height);
        const std::size_t n = static_cast<std::size_t>(w) * h;
        std::vector<float> lum(n);
        std::vector<double> weight(n), rowWeight(h, 0.0);
        const double pi = 3.14159265358979323846;
        double total = 0.0;
        for (std::uint32_t y = 0; y < h; ++y) {
            const double solidAngle = 2.0 * pi / static_cast<double>(w) *
                std::abs(std::cos(pi * (1.0 - static_cast<double>(y + 1U) / h)) - std::cos(pi * (1.0 - static_cast<double>(y) / h)));
            for (std::uint32_t x = 0; x < w; ++x) {
                const std::size_t p = static_cast<std::size_t>(y) * w + x;
                float rgb[3];
                LinearRgb(src, p, rgb);
                lum[p] = 0.2126F * rgb[0] + 0.7152F * rgb[1] + 0.0722F * rgb[2];
                weight[p] = static_cast<double>(lum[p]) * solidAngle;
                rowWeight[y] += weight[p];
                total += weight[p];
            }
        }
        if (!(total > 0.0 && std::isfinite(total))) return kMissing;
        Texture t;
        t.width = src.width;
        t.height = src.height;
        t.texels.assign(n * 4, 1.0F);
        double marginal = 0.0;
        for (std::uint32_t y = 0; y < h; ++y) {
            marginal += rowWeight[y] / total;
            double conditional = 0.0;
            for (std::uint32_t x = 0; x < w; ++x) {
                const std::size_t p = static_cast<std::size_t>(y) * w + x;
                conditional += rowWeight[y] > 0.0 ? weight[p] / rowWeight[y] : 1.0 / static_cast<double>(w);
                t.texels[p * 4] = x + 1U == w ? 1.0F : static_cast<float>(conditional);
                t.texels[p * 4 + 1] = y + 1U == h ? 1.0F : static_cast<float>(marginal);
                t.texels[p * 4 + 2] = static_cast<float>(static_cast<double>(lum[p]) / total);
            }
        }
        textures.push_back(std::move(t));
        return static_cast<std::uint32_t>(textures.size() - 1);
    }

    // Binned-SAH BVH2 (16 bins per axis over the centroid bounds, node and
    // triangle costs 1, leaves of at most 4 triangles; a range whose
    // centroids coincide is halved). Only a culling structure: which
    // triangle is hit is decided by the triangle test and the tie rule, so
    // any conservative BVH returns the same answer as brute force; SAH
    // makes the CPU traversal (and the timed FP32 baseline) faster.
    static constexpr int kBins = 16;
    static constexpr std::uint32_t kMaxDepth = 500;  // < the traversal stack of ForEachCandidate
    struct Box {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        void Grow(const float* l, const float* h)
        {
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], l[k]);
                hi[k] = std::max(hi[k], h[k]);
            }
        }
        double HalfArea() const
        {
            if (!(lo[0] <= hi[0])) return 0.0;
            const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
            return dx * dy + dy * dz + dz * dx;
        }
    };

    void BuildBvh()
    {
        order.resize(triangleCount);
        for (std::uint32_t i = 0; i < triangleCount; ++i) order[i] = i;
        if (triangleCount == 0) return;
        std::vector<float> centroid(static_cast<std::size_t>(triangleCount) * 3);
        std::vector<float> tlo(static_cast<std::size_t>(triangleCount) * 3), thi(tlo.size());
        for (std::uint32_t t = 0; t < triangleCount; ++t)
            for (int k = 0; k < 3; ++k) {
                const float a = positions[indices[t * 3] * 4 + k], b = positions[indices[t * 3 + 1] * 4 + k],
                            c = positions[indices[t * 3 + 2] * 4 + k];
                centroid[t * 3 + k] = (a + b + c) / 3.0F;
                tlo[t * 3 + k] = std::min(a, std::min(b, c));
                thi[t * 3 + k] = std::max(a, std::max(b, c));
            }
        nodes.reserve(2 * static_cast<std::size_t>(triangleCount) / 3 + 2);
        struct Item { std::uint32_t node, begin, end, depth; };
        std::vector<Item> work;
        nodes.push_back({});
        work.push_back({0, 0, triangleCount, 1});
        while (!work.empty()) {
            const Item it = work.back();
            if (it.depth > kMaxDepth) throw std::runtime_error("oracle BVH deeper than its traversal stack");
            work.pop_back();
            Box box, cbox;
            for (std::uint32_t i = it.begin; i < it.end; ++i) {
                const std::uint32_t t = order[i];
                box.Grow(&tlo[t * 3], &thi[t * 3]);
                cbox.Grow(&centroid[t * 3], &centroid[t * 3]);
            }
            BvhNode& n = nodes[it.node];
            for (int k = 0; k < 3; ++k) {
                const float pad = 1e-5F * std::max(std::fabs(box.lo[k]), std::fabs(box.hi[k])) + 1e-20F;
                n.lo[k] = box.lo[k] - pad;
                n.hi[k] = box.hi[k] + pad;
            }
            const std::uint32_t count = it.end - it.begin;
            // Best binned split over the three axes.
            int bestAxis = -1, bestBin = 0;
            double bestCost = INFINITY;
            for (int axis = 0; axis < 3; ++axis) {
                const float extent = cbox.hi[axis] - cbox.lo[axis];
                if (!(extent > 0.0F)) continue;
                Box bins[kBins];
                std::uint32_t counts[kBins] = {};
                const float scale = float(kBins) / extent;
                for (std::uint32_t i = it.begin; i < it.end; ++i) {
                    const std::uint32_t t = order[i];
                    const int b = std::min(kBins - 1, static_cast<int>((centroid[t * 3 + axis] - cbox.lo[axis]) * scale));
                    ++counts[b];
                    bins[b].Grow(&tlo[t * 3], &thi[t * 3]);
                }
                double rightArea[kBins];
                std::uint32_t rightCount[kBins];
                Box acc;
                std::uint32_t c = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.Grow(bins[b].lo, bins[b].hi);
                    c += counts[b];
                    rightArea[b] = acc.HalfArea();
                    rightCount[b] = c;
                }
                Box left;
                std::uint32_t lc = 0;
                for (int b = 1; b < kBins; ++b) {
                    left.Grow(bins[b - 1].lo, bins[b - 1].hi);
                    lc += counts[b - 1];
                    if (lc == 0 || rightCount[b] == 0) continue;
                    const double cost = left.HalfArea() * lc + rightArea[b] * rightCount[b];
                    if (cost < bestCost) {
                        bestCost = cost;
                        bestAxis = axis;
                        bestBin = b;
                    }
                }
            }
            // Leaf when testing the range costs no more than one more node visit
            // plus the children's expected tests (SAH, costs 1); at most 4.
            const double area = std::max(box.HalfArea(), 1e-30);
            if (count <= 1 || (count <= 4 && (bestAxis < 0 || 1.0 + bestCost / area >= double(count)))) {
                n.left = it.begin;
                n.count = count;
                n.right = 0;
                continue;
            }
            std::uint32_t mid;
            if (bestAxis >= 0) {
                const float lo = cbox.lo[bestAxis], scale = float(kBins) / (cbox.hi[bestAxis] - lo);
                const auto firstRight = std::partition(order.begin() + it.begin, order.begin() + it.end, [&](std::uint32_t t) {
                    return std::min(kBins - 1, static_cast<int>((centroid[t * 3 + bestAxis] - lo) * scale)) < bestBin;
                });
                mid = static_cast<std::uint32_t>(firstRight - order.begin());
            } else {
                mid = it.begin + count / 2;  // coincident centroids: any halving is conservative
            }
            const std::uint32_t l = static_cast<std::uint32_t>(nodes.size());
            nodes.push_back({});
            nodes.push_back({});
            BvhNode& nn = nodes[it.node];
            nn.left = l;
            nn.right = l + 1;
            nn.count = 0;
            work.push_back({l, it.begin, mid, it.depth + 1});
            work.push_back({l + 1, mid, it.end, it.depth + 1});
        }
    }
};

// ---- Watertight ray/triangle test (Woop, Benthin, Wald 2013) ----
#pragma GCC push_options
template <class R>
struct RayPre {
    V3<R> o, d;
    int kx, ky, kz;
    R sx, sy, sz;
};

template <class R>
RayPre<R> Prepare(const V3<R>& o, const V3<R>& d)
{
    RayPre<R> p;
    p.o = o;
    p.d = d;
    const R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0;
    R m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < R(0)) std::swap(kx, ky);
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    p.sx = d[kx] / d[kz];
    p.sy = d[ky] / d[kz];
    p.sz = R(1) / d[kz];
    return p;
}

// Returns true and (t, u, v) if the triangle is hit with tmin <= t <= tmax.
// u, v are the barycentrics of vertices 1 and 2.
template <class R>
bool IntersectTriangle(const RayPre<R>& r, const V3<R>& p0, const V3<R>& p1, const V3<R>& p2,
                       R tmin, R tmax, R& t, R& u, R& v)
{
    const V3<R> a = p0 - r.o, b = p1 - r.o, c = p2 - r.o;
    const R az0 = a[r.kz], bz0 = b[r.kz], cz0 = c[r.kz];
    const R ax = a[r.kx] - r.sx * az0, ay = a[r.ky] - r.sy * az0;
    const R bx = b[r.kx] - r.sx * bz0, by = b[r.ky] - r.sy * bz0;
    const R cx = c[r.kx] - r.sx * cz0, cy = c[r.ky] - r.sy * cz0;
    R U = cx * by - cy * bx;
    R V = ax * cy - ay * cx;
    R W = bx * ay - by * ax;
    if (sizeof(R) == 4 && (U == R(0) || V == R(0) || W == R(0))) {
        const double cxd = cx, cyd = cy, bxd = bx, byd = by, axd = ax, ayd = ay;
        U = static_cast<R>(cxd * byd - cyd * bxd);
        V = static_cast<R>(axd * cyd - ayd * cxd);
        W = static_cast<R>(bxd * ayd - byd * axd);
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    const R det = U + V + W;
    if (det == R(0)) return false;
    const R az = r.sz * az0, bz = r.sz * bz0, cz = r.sz * cz0;
    const R T = U * az + V * bz + W * cz;
    t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    u = V / det;
    v = W / det;
    return true;
}
#pragma GCC pop_options

bool BoxHit(const BvhNode& n, const double o[3], const double inv[3], double tmin, double tmax)
{
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (n.lo[k] - o[k]) * inv[k];
        double b = (n.hi[k] - o[k]) * inv[k];
        if (std::isnan(a)) a = -INFINITY;  // 0 * inf on an axis-parallel ray
        if (std::isnan(b)) b = INFINITY;
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b * (1.0 + 1e-9));
    }
    return t0 <= t1;
}

// Visits every triangle whose BVH leaf overlaps [tmin, tmax]; visit(prim)
// returns false to stop.
template <class R, class Visit>
void ForEachCandidate(const Scene& s, const V3<R>& o, const V3<R>& d, R tmin, R tmax, Visit&& visit)
{
    if (s.nodes.empty()) return;
    const double od[3] = {double(o.x), double(o.y), double(o.z)};
    const double inv[3] = {1.0 / double(d.x), 1.0 / double(d.y), 1.0 / double(d.z)};
    std::uint32_t stack[512];  // the build rejects trees deeper than 500 levels
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode& n = s.nodes[stack[--sp]];
        if (!BoxHit(n, od, inv, double(tmin) - 1e-6, double(tmax) * (1.0 + 1e-6) + 1e-6)) continue;
        if (n.count > 0) {
            for (std::uint32_t i = 0; i < n.count; ++i)
                if (!visit(s.order[n.left + i])) return;
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
}

template <class R>
bool TestPrimitive(const Scene& s, const RayPre<R>& pre, std::uint32_t prim, R tmin, R tmax, R& t, R& u, R& v)
{
    return IntersectTriangle(pre, s.Pos<R>(s.indices[prim * 3]), s.Pos<R>(s.indices[prim * 3 + 1]),
                             s.Pos<R>(s.indices[prim * 3 + 2]), tmin, tmax, t, u, v);
}

// Closest hit among all triangles (ignores opacity).
template <class R>
Hit ClosestHit(const Scene& s, const V3<R>& o, const V3<R>& d, R tmin, R tmax)
{
    Hit best;
    const RayPre<R> pre = Prepare(o, d);
    R bestT = tmax;
    ForEachCandidate<R>(s, o, d, tmin, tmax, [&](std::uint32_t prim) {
        R t, u, v;
        if (TestPrimitive(s, pre, prim, tmin, bestT, t, u, v)) {
            if (!best.found || t < R(best.t) || (t == R(best.t) && prim < best.primitive)) {
                best.found = true;
                best.primitive = prim;
                best.t = double(t);
                best.u = double(u);
                best.v = double(v);
                bestT = t;
            }
        }
        return true;
    });
    return best;
}

template <class R>
Hit ClosestHitBruteForce(const Scene& s, const V3<R>& o, const V3<R>& d, R tmin, R tmax)
{
    Hit best;
    const RayPre<R> pre = Prepare(o, d);
    for (std::uint32_t prim = 0; prim < s.triangleCount; ++prim) {
        R t, u, v;
        if (TestPrimitive(s, pre, prim, tmin, tmax, t, u, v)) {
            if (!best.found || t < R(best.t) || (t == R(best.t) && prim < best.primitive)) {
                best.found = true;
                best.primitive = prim;
                best.t = double(t);
                best.u = double(u);
                best.v = double(v);
            }
        }
    }
    return best;
}

// All hits in [tmin, tmax] sorted by (t, primitive).
template <class R>
void AllHits(const Scene& s, const V3<R>& o, const V3<R>& d, R tmin, R tmax, std::vector<Hit>& out)
{
    out.clear();
    const RayPre<R> pre = Prepare(o, d);
    ForEachCandidate<R>(s, o, d, tmin, tmax, [&](std::uint32_t prim) {
        R t, u, v;
        if (TestPrimitive(s, pre, prim, tmin, tmax, t, u, v)) {
            Hit h;
            h.found = true;
            h.primitive = prim;
            h.t = double(t);
            h.u = double(u);
            h.v = double(v);
            out.push_back(h);
        }
        return true;
    });
    std::sort(out.begin(), out.end(), [](const Hit& a, const Hit& b) {
        return a.t < b.t || (a.t == b.t && a.primitive < b.primitive);
    });
}

// ======================= Shading restatement ============================

template <class R>
struct Integrator {
    const Scene& s;
    std::uint32_t lightCount;
    bool zUp;
    bool opaqueFlag;

    // reconstructRgb (:118-126)
    static R ReconstructRgb(const V3<R>& rgb, R w)
    {
        const V3<R> basis(std::exp(R(-0.5) * sq((w - R(620)) / R(30))),
                          std::exp(R(-0.5) * sq((w - R(535)) / R(25))),
                          std::exp(R(-0.5) * sq((w - R(450)) / R(20))));
        return dot(vmax(rgb, R(0)), basis) / std::max(dot(V3<R>(R(1)), basis), R(1e-5));
    }
    // reflectanceSpectrum (:128-131)
    static R Reflectance(const V3<R>& rgb, R w) { return clampR(ReconstructRgb(rgb, w), R(0), R(1)); }
    // illuminantSpectrum (:133-144)
    static R Illuminant(const V3<R>& rgb, R w)
    {
        const R temperature = R(6504), c2 = R(1.4387769e7), ref = R(560);
        const R planck = std::pow(ref / w, R(5)) * (std::exp(c2 / (ref * temperature)) - R(1)) /
                         std::max(std::exp(c2 / (w * temperature)) - R(1), R(1e-5));
        return std::max(ReconstructRgb(rgb, w) * planck, R(0));
    }
    // cieXyz (:146-170): multi-lobe Gaussian fit of Wyman, Sloan, Shirley 2013.
    static V3<R> CieXyz(R w)
    {
        const R tx1 = (w - R(442.0)) * (w < R(442.0) ? R(0.0624) : R(0.0374));
        const R tx2 = (w - R(599.8)) * (w < R(599.8) ? R(0.0264) : R(0.0323));
        const R tx3 = (w - R(501.1)) * (w < R(501.1) ? R(0.0490) : R(0.0382));
        const R ty1 = (w - R(568.8)) * (w < R(568.8) ? R(0.0213) : R(0.0247));
        const R ty2 = (w - R(530.9)) * (w < R(530.9) ? R(0.0613) : R(0.0322));
        const R tz1 = (w - R(437.0)) * (w < R(437.0) ? R(0.0845) : R(0.0278));
        const R tz2 = (w - R(459.0)) * (w < R(459.0) ? R(0.0385) : R(0.0725));
        return {R(0.362) * std::exp(R(-0.5) * tx1 * tx1) + R(1.056) * std::exp(R(-0.5) * tx2 * tx2) -
                    R(0.065) * std::exp(R(-0.5) * tx3 * tx3),
                R(0.821) * std::exp(R(-0.5) * ty1 * ty1) + R(0.286) * std::exp(R(-0.5) * ty2 * ty2),
                R(1.217) * std::exp(R(-0.5) * tz1 * tz1) + R(0.681) * std::exp(R(-0.5) * tz2 * tz2)};
    }
    // spectralSensor (:172-179): XYZ scaled by 400/106.856895, then the
    // IEC 61966-2-1 XYZ -> linear sRGB matrix.
    static V3<R> Sensor(R w)
    {
        const V3<R> xyz = CieXyz(w) * (R(400) / R(106.856895));
        return {R(3.2404542) * xyz.x - R(1.5371385) * xyz.y - R(0.4985314) * xyz.z,
                R(-0.9692660) * xyz.x + R(1.8760108) * xyz.y + R(0.0415560) * xyz.z,
                R(0.0556434) * xyz.x - R(0.2040259) * xyz.y + R(1.0572252) * xyz.z};
    }

    static R Rand(std::uint32_t& st) { return R(RandomFloat(st)); }
    static R Pi() { return R(kPiD); }
    static R TwoPi() { return R(6.28318530718); }

    // Basis-completing tangent used by several samplers (:188-190).
    static V3<R> AnyTangent(const V3<R>& n)
    {
        return normalize(std::fabs(n.z) < R(0.999) ? cross(n, V3<R>(0, 0, 1)) : cross(n, V3<R>(0, 1, 0)));
    }
    // cosineHemisphere (:183-195)
    static V3<R> CosineHemisphere(const V3<R>& n, std::uint32_t& st)
    {
        const R r1 = TwoPi() * Rand(st);
        const R r2 = Rand(st);
        const R r2s = std::sqrt(r2);
        const V3<R> t = AnyTangent(n);
        const V3<R> b = cross(n, t);
        return normalize(t * (std::cos(r1) * r2s) + b * (std::sin(r1) * r2s) + n * std::sqrt(R(1) - r2));
    }
    // luminance (:197-200) of a grey spectral vec3 s.
    static R LumGrey(R s) { return s * R(0.2126) + s * R(0.7152) + s * R(0.0722); }
    // fresnelSchlick (:202-205)
    static R Fresnel(R c, R f0) { return f0 + (R(1) - f0) * std::pow(R(1) - clampR(c, R(0), R(1)), R(5)); }
    // orthogonalTangent (:276-285)
    static V3<R> OrthoTangent(const V3<R>& n, const V3<R>& t)
    {
        V3<R> r = t - n * dot(n, t);
        if (dot(r, r) < R(1e-8)) r = std::fabs(n.z) < R(0.999) ? cross(n, V3<R>(0, 0, 1)) : cross(n, V3<R>(0, 1, 0));
        return normalize(r);
    }
    struct Alpha { R x, y; };
    static Alpha AlphaOf(R rx, R ry)
    {
        return {clampR(rx * rx, R(0.001), R(1)), clampR(ry * ry, R(0.001), R(1))};
    }
    // distributionGGXAnisotropic (:287-300)
    static R DAniso(const V3<R>& h, const V3<R>& n, const V3<R>& t, R rx, R ry)
    {
        const V3<R> axX = OrthoTangent(n, t);
        const V3<R> axY = cross(n, axX);
        const Alpha a = AlphaOf(rx, ry);
        const R lx = dot(h, axX), ly = dot(h, axY), lz = std::max(dot(h, n), R(0));
        const R den = lx * lx / (a.x * a.x) + ly * ly / (a.y * a.y) + lz * lz;
        return R(1) / std::max(Pi() * a.x * a.y * den * den, R(1e-6));
    }
    // geometryGGXAnisotropic (:302-315)
    static R GAniso(const V3<R>& dir, const V3<R>& n, const V3<R>& t, R rx, R ry)
    {
        const R nd = std::max(dot(n, dir), R(0));
        if (nd <= R(0)) return R(0);
        const V3<R> axX = OrthoTangent(n, t);
        const V3<R> axY = cross(n, axX);
        const Alpha a = AlphaOf(rx, ry);
        const R proj = sq(a.x * dot(dir, axX)) + sq(a.y * dot(dir, axY));
        const R lambda = (R(-1) + std::sqrt(R(1) + proj / std::max(nd * nd, R(1e-8)))) * R(0.5);
        return R(1) / (R(1) + lambda);
    }
    // evaluateGGXAnisotropic (:317-332)
    static R EvalGGX(const V3<R>& n, const V3<R>& t, const V3<R>& v, const V3<R>& l, R rx, R ry, R f0)
    {
        const V3<R> h = normalize(v + l);
        const R nv = std::max(dot(n, v), R(0)), nl = std::max(dot(n, l), R(0)), vh = std::max(dot(v, h), R(0));
        if (nv <= R(0) || nl <= R(0) || vh <= R(0)) return R(0);
        const R D = DAniso(h, n, t, rx, ry);
        const R G = GAniso(v, n, t, rx, ry) * GAniso(l, n, t, rx, ry);
        return D * G * Fresnel(vh, f0) / std::max(R(4) * nv * nl, R(1e-5));
    }
    // sampleGGXReflectionAnisotropic (:334-350)
    static V3<R> SampleGGX(const V3<R>& inc, const V3<R>& n, const V3<R>& t, R rx, R ry, std::uint32_t& st)
    {
        const V3<R> axX = OrthoTangent(n, t);
        const V3<R> axY = cross(n, axX);
        const Alpha a = AlphaOf(rx, ry);
        const R phi = TwoPi() * Rand(st);
        const R e = Rand(st);
        const R radius = std::sqrt(e / std::max(R(1) - e, R(1e-6)));
        V3<R> h = normalize(axX * (a.x * radius * std::cos(phi)) + axY * (a.y * radius * std::sin(phi)) + n);
        if (dot(h, -inc) < R(0)) h = -h;
        const V3<R> r = reflect(inc, h);
        return dot(r, n) > R(0) ? normalize(r) : normalize(reflect(inc, n));
    }
    // ggxSampleWeightAnisotropic (:352-367)
    static R GGXWeight(const V3<R>& inc, const V3<R>& out, const V3<R>& n, const V3<R>& t, R rx, R ry, R f0)
    {
        const V3<R> v = normalize(-inc);
        const V3<R> h = normalize(v + out);
        const R nl = std::max(dot(n, out), R(0)), nh = std::max(dot(n, h), R(0)), vh = std::max(dot(v, h), R(0));
        if (nl <= R(0) || nh <= R(0) || vh <= R(0)) return R(0);
        const R D = DAniso(h, n, t, rx, ry);
        const R pdf = D * nh / std::max(R(4) * vh, R(1e-6));
        return EvalGGX(n, t, v, out, rx, ry, f0) * nl / std::max(pdf, R(1e-6));
    }
    // ggxReflectionPdfAnisotropic (:371-383)
    static R GGXPdf(const V3<R>& n, const V3<R>& t, const V3<R>& v, const V3<R>& l, R rx, R ry)
    {
        if (dot(n, v) <= R(0) || dot(n, l) <= R(0)) return R(0);
        const V3<R> h = normalize(v + l);
        const R nh = std::max(dot(n, h), R(0)), vh = std::max(dot(v, h), R(0));
        return DAniso(h, n, t, rx, ry) * nh / std::max(R(4) * vh, R(1e-6));
    }
    // powerHeuristic (:385-390)
    static R Power(R a, R b)
    {
        const R a2 = a * a, b2 = b * b;
        return a2 / std::max(a2 + b2, R(1e-12));
    }
    // surfaceReflectionPdf (:392-415)
    static R SurfacePdf(const V3<R>& n, const V3<R>& t, const V3<R>& v, const V3<R>& l, R rx, R ry, R crx, R cry,
                        R pSheen, R pCoat, R pSpec, R pTrans)
    {
        const R refl = std::max(dot(n, l), R(0)) / Pi();
        const R trans = std::max(-dot(n, l), R(0)) / Pi();
        const R coatPdf = GGXPdf(n, t, v, l, crx, cry);
        const R specPdf = GGXPdf(n, t, v, l, rx, ry);
        const R diffPdf = mixR(refl, trans, pTrans);
        const R base = pCoat * coatPdf + (R(1) - pCoat) * (pSpec * specPdf + (R(1) - pSpec) * diffPdf);
        return pSheen * refl + (R(1) - pSheen) * base;
    }
    // sheenDirectionalAlbedo (:419-429)
    static R SheenAlbedoCK(R nv, R rough)
    {
        const R x = clampR(nv, R(0), R(1)), y = clampR(rough, R(0.01), R(1));
        const R fx = R(13.67300) + R(-68.78018) * x + R(799.08825) * y + R(-905.00061) * x * y +
                     R(60.28956) * x * x + R(1086.96473) * y * y;
        const R fy = R(1.0) + R(61.57746) * x + R(442.78211) * y + R(2597.49308) * x * y +
                     R(121.81241) * x * x + R(3045.55075) * y * y;
        return clampR(fx / std::max(fy, R(1e-6)), R(0), R(1));
    }
    // sheenImageworksResponse (:431-446)
    static R SheenImageworks(const V3<R>& n, const V3<R>& v, const V3<R>& l, R rough)
    {
        const R nl = std::max(dot(n, l), R(0)), nv = std::max(dot(n, v), R(0));
        if (nl <= R(0) || nv <= R(0)) return R(0);
        const V3<R> h = normalize(v + l);
        const R nh = std::max(dot(n, h), R(0));
        const R inv = R(1) / std::max(rough, R(0.005));
        const R D = (R(2) + inv) * std::pow(std::max(R(1) - nh * nh, R(0)), inv * R(0.5)) / (R(2) * Pi());
        const R brdf = D / std::max(R(4) * (nl + nv - nl * nv), R(1e-6));
        return brdf * nl;
    }
    // sheenZeltnerDirectionalAlbedo (:448-460)
    static R SheenAlbedoZ(R nv, R rough)
    {
        const R x = clampR(nv, R(0), R(1)), y = clampR(rough, R(0.01), R(1));
        const R sd = y * (R(0.0206607) + R(1.58491) * y) / (R(0.0379424) + y * (R(1.32227) + y));
        const R m = y * (R(-0.193854) + y * (R(-1.14885) + y * (R(1.7932) - R(0.95943) * y * y))) / (R(0.046391) + y);
        const R o = y * (R(0.000654023) + (R(-0.0207818) + R(0.119681) * y) * y) / (R(1.26264) + y * (R(-1.92021) + y));
        return clampR(std::exp(R(-0.5) * sq((x - m) / std::max(sd, R(1e-6)))) /
                          (std::max(sd, R(1e-6)) * std::sqrt(R(2) * Pi())) + o,
                      R(0), R(1));
    }
    // sheenResponse (:462-471)
    static R SheenResponse(const V3<R>& n, const V3<R>& v, const V3<R>& l, R rough, bool z)
    {
        if (!z) return SheenImageworks(n, v, l, rough);
        const R nl = std::max(dot(n, l), R(0)), nv = std::max(dot(n, v), R(0));
        return SheenAlbedoZ(nv, rough) * nl / Pi();
    }
    // directionToVerticalCross (:473-503)
    static void VerticalCross(const V3<R>& d, R& u, R& v)
    {
        const V3<R> a(std::fabs(d.x), std::fabs(d.y), std::fabs(d.z));
        R fu, fv, tu, tv;
        if (a.x >= a.y && a.x >= a.z) {
            if (d.x >= R(0)) { fu = -d.z / a.x; fv = d.y / a.x; tu = 2; tv = 1; }
            else { fu = d.z / a.x; fv = d.y / a.x; tu = 0; tv = 1; }
        } else if (a.y >= a.z) {
            if (d.y >= R(0)) { fu = d.x / a.y; fv = -d.z / a.y; tu = 1; tv = 0; }
            else { fu = d.x / a.y; fv = d.z / a.y; tu = 1; tv = 2; }
        } else if (d.z >= R(0)) { fu = d.x / a.z; fv = d.y / a.z; tu = 1; tv = 1; }
        else { fu = -d.x / a.z; fv = d.y / a.z; tu = 1; tv = 3; }
        u = (tu + fu * R(0.5) + R(0.5)) / R(3);
        v = (tv + fv * R(0.5) + R(0.5)) / R(4);
    }
    // domeTextureCoordinates (:505-525)
    static void DomeUv(const V3<R>& d, std::uint32_t format, R& u, R& v)
    {
        if (format == 2U) {
            const R den = R(2) * std::sqrt(std::max(d.x * d.x + d.y * d.y + (d.z + R(1)) * (d.z + R(1)), R(1e-8)));
            u = d.x / den + R(0.5);
            v = d.y / den + R(0.5);
            return;
        }
        if (format == 3U) {
            const R radius = std::acos(clampR(-d.z, R(-1), R(1))) / Pi();
            const R len = std::sqrt(d.x * d.x + d.y * d.y);
            R rx = 0, ry = 0;
            if (len > R(1e-6)) { rx = d.x / len; ry = d.y / len; }
            u = R(0.5) + rx * (R(0.5) * radius);
            v = R(0.5) + ry * (R(0.5) * radius);
            return;
        }
        if (format == 4U) { VerticalCross(d, u, v); return; }
        u = fractR(std::atan2(d.x, -d.z) / (R(2) * Pi()) + R(1));
        v = R(1) - std::acos(clampR(d.y, R(-1), R(1))) / Pi();
    }

    // Bilinear REPEAT filtering of a decoded texture (sampler :2401-2423).
    V3<R> TexRgb(std::uint32_t index, R u, R v, R* alpha = nullptr) const
    {
        const Texture& t = s.textures[index];
        const R x = u * R(t.width) - R(0.5), y = v * R(t.height) - R(0.5);
        const R fx = std::floor(x), fy = std::floor(y);
        const R ax = x - fx, ay = y - fy;
        auto wrap = [](long long i, int n) { long long m = i % n; return static_cast<int>(m < 0 ? m + n : m); };
        long long ix = static_cast<long long>(fx), iy = static_cast<long long>(fy);
        if (!std::isfinite(static_cast<double>(x)) || !std::isfinite(static_cast<double>(y))) { ix = 0; iy = 0; }
        const int x0 = wrap(ix, t.width), x1 = wrap(ix + 1, t.width);
        const int y0 = wrap(iy, t.height), y1 = wrap(iy + 1, t.height);
        R out[4];
        for (int c = 0; c < 4; ++c) {
            const R t00 = R(t.texels[(static_cast<std::size_t>(y0) * t.width + x0) * 4 + c]);
            const R t10 = R(t.texels[(static_cast<std::size_t>(y0) * t.width + x1) * 4 + c]);
            const R t01 = R(t.texels[(static_cast<std::size_t>(y1) * t.width + x0) * 4 + c]);
            const R t11 = R(t.texels[(static_cast<std::size_t>(y1) * t.width + x1) * 4 + c]);
            out[c] = (t00 * (R(1) - ax) + t10 * ax) * (R(1) - ay) + (t01 * (R(1) - ax) + t11 * ax) * ay;
        }
        if (alpha) *alpha = out[3];
        return {out[0], out[1], out[2]};
    }
    float Texel(std::uint32_t index, int x, int y, int c) const
    {
        const Texture& t = s.textures[index];
        return t.texels[(static_cast<std::size_t>(y) * t.width + x) * 4 + c];
    }

    // sampleMaterialTexture (:760-775): returns RGBA.
    void SampleMat(std::uint32_t ti, R u, R v, const R fallback[4], R out[4]) const
    {
        if (ti == kMissing) { for (int c = 0; c < 4; ++c) out[c] = fallback[c]; return; }
        std::uint32_t texIndex = ti;
        R su = u, sv = v;
        if ((ti & 0x80000000U) != 0U) {
            const R tu = std::floor(u), tv = std::floor(v);
            const long long tx = static_cast<long long>(tu), ty = static_cast<long long>(tv);
            const long long off = tx + ty * 10;
            if (off < 0 || off >= 23) { for (int c = 0; c < 4; ++c) out[c] = fallback[c]; return; }
            const std::uint32_t present = (ti >> 8U) & 0x7fffffU;
            const std::uint32_t bit = 1U << static_cast<std::uint32_t>(off);
            if ((present & bit) == 0U) { for (int c = 0; c < 4; ++c) out[c] = fallback[c]; return; }
            texIndex = (ti & 0xffU) + static_cast<std::uint32_t>(__builtin_popcount(present & (bit - 1U)));
            su = fractR(u);
            sv = fractR(v);
        }
        R a;
        const V3<R> rgb = TexRgb(texIndex, su, sv, &a);
        out[0] = rgb.x; out[1] = rgb.y; out[2] = rgb.z; out[3] = a;
    }
    R SampleMatR(std::uint32_t ti, R u, R v, R fb) const
    {
        const R f[4] = {fb, fb, fb, fb};
        R o[4];
        SampleMat(ti, u, v, f, o);
        return o[0];
    }
    V3<R> SampleMatRgb(std::uint32_t ti, R u, R v, const V3<R>& fb) const
    {
        const R f[4] = {fb.x, fb.y, fb.z, R(1)};
        R o[4];
        SampleMat(ti, u, v, f, o);
        return {o[0], o[1], o[2]};
    }

    static V3<R> XYZ(const F4& f) { return {R(f.x), R(f.y), R(f.z)}; }

    // sampleDome (:527-537)
    V3<R> SampleDome(const Light& L, const V3<R>& wd) const
    {
        const V3<R> local = normalize(V3<R>(dot(wd, XYZ(L.basisX)), dot(wd, XYZ(L.basisY)), dot(wd, XYZ(L.basisZ))));
        V3<R> tex(R(1));
        if (L.textureInfo.x != kMissing) {
            R u, v;
            DomeUv(local, L.textureInfo.y, u, v);
            tex = TexRgb(L.textureInfo.x, u, v);
        }
        return XYZ(L.colorIntensity) * R(L.colorIntensity.w) * tex;
    }
    // hasEmissiveTriangles (:539-549)
    bool HasEmissive() const
    {
        const std::uint32_t n = s.triangleCount;
        return n > 0U && s.emissive[(n - 1U) * 4] > 0.0F;
    }
    // environment (:551-569)
    V3<R> Environment(const V3<R>& d) const
    {
        V3<R> result(R(0));
        if (lightCount == 0U) {
            if (HasEmissive()) return V3<R>(R(0));
            const V3<R> up = zUp ? V3<R>(0, 0, 1) : V3<R>(0, 1, 0);
            const R t = R(0.5) * (dot(d, up) + R(1));
            return V3<R>(R(0.035), R(0.045), R(0.065)) * (R(1) - t) + V3<R>(R(0.38), R(0.55), R(0.85)) * t;
        }
        for (std::uint32_t i = 0; i < lightCount; ++i)
            if (s.lights[i].textureInfo.z == 0U) result += SampleDome(s.lights[i], d);
        return result;
    }
    // latLongDirection (:571-579)
    static V3<R> LatLongDirection(R u, R v)
    {
        const R theta = Pi() * (R(1) - v), phi = R(2) * Pi() * u;
        const R st = std::sin(theta);
        return {st * std::sin(phi), std::cos(theta), -st * std::cos(phi)};
    }
    // uniformSphere (:1077-1083)
    static V3<R> UniformSphere(std::uint32_t& st)
    {
        const R y = R(1) - R(2) * Rand(st);
        const R radius = std::sqrt(std::max(R(1) - y * y, R(0)));
        const R phi = TwoPi() * Rand(st);
        return {radius * std::cos(phi), y, radius * std::sin(phi)};
    }
    // domeLightPdf (:583-598)
    R DomePdf(const Light& L, const V3<R>& wd) const
    {
        if (L.textureInfo.w == kMissing) return R(1) / R(12.5663706144);
        const V3<R> local = normalize(V3<R>(dot(wd, XYZ(L.basisX)), dot(wd, XYZ(L.basisY)), dot(wd, XYZ(L.basisZ))));
        R u, v;
        DomeUv(local, L.textureInfo.y, u, v);
        const Texture& imp = s.textures[L.textureInfo.w];
        const int tx = std::clamp(static_cast<int>(u * R(imp.width)), 0, imp.width - 1);
        const int ty = std::clamp(static_cast<int>(v * R(imp.height)), 0, imp.height - 1);
        return std::max(R(Texel(L.textureInfo.w, tx, ty, 2)), R(0));
    }
    // sampleDomeLight (:600-650)
    void SampleDomeLight(const Light& L, std::uint32_t& rng, V3<R>& wd, R& pdf) const
    {
        if (L.textureInfo.w == kMissing) {
            wd = UniformSphere(rng);
            pdf = R(1) / R(12.5663706144);
            return;
        }
        const std::uint32_t ii = L.textureInfo.w;
        const Texture& imp = s.textures[ii];
        const R rowTarget = Rand(rng);
        int lo = 0, hi = imp.height - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (rowTarget <= R(Texel(ii, 0, mid, 1))) hi = mid; else lo = mid + 1;
        }
        const int row = lo;
        const R colTarget = Rand(rng);
        lo = 0;
        hi = imp.width - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (colTarget <= R(Texel(ii, mid, row, 0))) hi = mid; else lo = mid + 1;
        }
        const int col = lo;
        const R v0 = R(row) / R(imp.height), v1 = R(row + 1) / R(imp.height);
        const R c0 = std::cos(Pi() * (R(1) - v0)), c1 = std::cos(Pi() * (R(1) - v1));
        const R ct = mixR(c0, c1, Rand(rng));
        const R v = R(1) - std::acos(clampR(ct, R(-1), R(1))) / Pi();
        const R u = (R(col) + Rand(rng)) / R(imp.width);
        const V3<R> local = LatLongDirection(u, v);
        wd = normalize(XYZ(L.basisX) * local.x + XYZ(L.basisY) * local.y + XYZ(L.basisZ) * local.z);
        pdf = std::max(R(Texel(ii, col, row, 2)), R(0));
    }
    // lightSelectionWeight (:652-691)
    static R SelectionWeight(const Light& L, const V3<R>& pos, const V3<R>& n, R de, R se, R te)
    {
        const R power = std::max(R(L.geometry.w), R(0));
        if (power <= R(0)) return R(0);
        const R dr = std::max(de, R(0)) * R(L.controls.x);
        const R sr = std::max(se, R(0)) * R(L.controls.y);
        const R tr = std::max(te, R(0)) * R(L.controls.x);
        const R minimum = power * R(0.01);
        const std::uint32_t type = L.textureInfo.z;
        if (type == 0U) return std::max(minimum, power * (R(0.5) * (dr + tr) + sr));
        V3<R> dir;
        R geom = R(1);
        if (type == 5U) {
            dir = normalize(XYZ(L.basisZ));
        } else {
            const V3<R> toL = XYZ(L.position) - pos;
            const R d2 = std::max(dot(toL, toL), R(1e-4));
            dir = toL * (R(1) / std::sqrt(d2));
            geom = R(1) / d2;
            if (type == 1U || type == 2U) {
                const V3<R> ln = -normalize(XYZ(L.basisZ));
                geom *= std::max(dot(ln, -dir), R(0));
            } else if (type == 4U) {
                geom *= R(0.5);
            }
        }
        const R nl = dot(n, dir);
        const R response = std::max(nl, R(0)) * (dr + sr) + std::max(-nl, R(0)) * tr;
        return std::max(minimum, power * geom * response);
    }
    // totalAnalyticLightWeight (:693-705)
    R TotalWeight(const V3<R>& pos, const V3<R>& n, R de, R se, R te) const
    {
        R total = R(0);
        for (std::uint32_t i = 0; i < lightCount; ++i) total += SelectionWeight(s.lights[i], pos, n, de, se, te);
        return total;
    }
    // sampleAnalyticLight (:707-737)
    void SampleAnalyticLight(const V3<R>& pos, const V3<R>& n, R de, R se, R te, std::uint32_t& rng,
                             std::uint32_t& index, R& pdf) const
    {
        const R total = TotalWeight(pos, n, de, se, te);
        if (total <= R(0)) {
            index = std::min(static_cast<std::uint32_t>(Rand(rng) * R(lightCount)), lightCount - 1U);
            pdf = R(1) / R(lightCount);
            return;
        }
        const R target = Rand(rng) * total;
        R acc = R(0);
        index = lightCount - 1U;
        for (std::uint32_t i = 0; i < lightCount; ++i) {
            acc += SelectionWeight(s.lights[i], pos, n, de, se, te);
            if (target <= acc) { index = i; break; }
        }
        pdf = SelectionWeight(s.lights[index], pos, n, de, se, te) / total;
    }
    // environmentLightPdf (:739-758)
    R EnvLightPdf(const V3<R>& pos, const V3<R>& n, const V3<R>& dir, R de, R se, R te) const
    {
        if (lightCount == 0U) return R(0);
        const R total = TotalWeight(pos, n, de, se, te);
        if (total <= R(0)) return R(0);
        R pdf = R(0);
        for (std::uint32_t i = 0; i < lightCount; ++i) {
            if (s.lights[i].textureInfo.z == 0U) {
                const R sel = SelectionWeight(s.lights[i], pos, n, de, se, te) / total;
                pdf += sel * DomePdf(s.lights[i], dir);
            }
        }
        return pdf;
    }
    const Material& MaterialOf(std::uint32_t prim) const { return s.materials[s.triangleMaterials[prim]]; }
    // surfaceTextureCoordinates (:787-795)
    void SurfaceUv(std::uint32_t prim, R bu, R bv, R& u, R& v) const
    {
        const R w0 = R(1) - bu - bv;
        const float* t = &s.texcoords[prim * 6];
        u = R(t[0]) * w0 + R(t[2]) * bu + R(t[4]) * bv;
        v = R(t[1]) * w0 + R(t[3]) * bu + R(t[5]) * bv;
    }
    // surfaceEmissionRgb (:777-785)
    V3<R> EmissionRgb(std::uint32_t prim, R u, R v) const
    {
        const Material& m = MaterialOf(prim);
        const R weight = std::max(R(m.transmissionOpacityIor.w), R(0));
        if (m.textureIndices0.w == kMissing) return XYZ(m.emissionRoughness) * weight;
        const R f[4] = {0, 0, 0, 0};
        R o[4];
        SampleMat(m.textureIndices0.w, u, v, f, o);
        return V3<R>(o[0], o[1], o[2]) * weight;
    }
    // surfaceOpacity (:797-803)
    R Opacity(std::uint32_t prim, R bu, R bv) const
    {
        const Material& m = MaterialOf(prim);
        R u, v;
        SurfaceUv(prim, bu, bv, u, v);
        return clampR(SampleMatR(m.textureIndices1.x, u, v, R(m.transmissionOpacityIor.y)), R(0), R(1));
    }
    // applyNormalMap (:807-829)
    V3<R> NormalMap(std::uint32_t ti, bool flipY, R scale, R u, R v, const V3<R>& gn, const V3<R>& p0,
                    const V3<R>& p1, const V3<R>& p2, R u0, R v0, R u1, R v1, R u2, R v2) const
    {
        if (ti == kMissing) return gn;
        const R du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
        const R det = du1 * dv2 - dv1 * du2;
        if (std::fabs(det) < R(1e-8)) return gn;
        const V3<R> e1 = p1 - p0, e2 = p2 - p0;
        V3<R> t = normalize((e1 * dv2 - e2 * dv1) / det);
        t = normalize(t - gn * dot(gn, t));
        V3<R> b = normalize(cross(gn, t));
        if (det < R(0)) b = -b;
        const R fb[4] = {R(0.5), R(0.5), R(1), R(1)};
        R o[4];
        SampleMat(ti, u, v, fb, o);
        V3<R> m(o[0] * R(2) - R(1), o[1] * R(2) - R(1), o[2] * R(2) - R(1));
        if (flipY) m.y = -m.y;
        m.x *= scale;
        m.y *= scale;
        const V3<R> r = normalize(t * m.x + b * m.y + gn * m.z);
        return dot(r, gn) < R(0) ? -r : r;
    }
    // surfaceTangent (:831-841)
    static V3<R> SurfaceTangent(const V3<R>& n, const V3<R>& p0, const V3<R>& p1, const V3<R>& p2, R u0, R v0,
                                R u1, R v1, R u2, R v2)
    {
        const R du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
        const R det = du1 * dv2 - dv1 * du2;
        if (std::fabs(det) < R(1e-8)) return OrthoTangent(n, V3<R>(1, 0, 0));
        const V3<R> t = ((p1 - p0) * dv2 - (p2 - p0) * dv1) / det;
        return OrthoTangent(n, t);
    }
    // surfaceShadingNormal (:843-853)
    V3<R> ShadingNormal(std::uint32_t prim, R bu, R bv, const V3<R>& gn) const
    {
        const R w0 = R(1) - bu - bv;
        const float* nn = &s.normals[prim * 12];
        V3<R> r = V3<R>(R(nn[0]), R(nn[1]), R(nn[2])) * w0 + V3<R>(R(nn[4]), R(nn[5]), R(nn[6])) * bu +
                  V3<R>(R(nn[8]), R(nn[9]), R(nn[10])) * bv;
        if (dot(r, r) < R(1e-12)) return gn;
        r = normalize(r);
        return dot(r, gn) < R(0) ? -r : r;
    }
    // Occlusion test: opaque scenes confirm any triangle; otherwise only
    // triangles of opacity >= 0.5; hitDistance is the closest confirmed.
    bool Occluded(const V3<R>& o, const V3<R>& d, R tmax, R& hitDistance) const
    {
        if (opaqueFlag) {
            const Hit h = ClosestHit<R>(s, o, d, R(0.001), tmax);
            if (h.found) hitDistance = R(h.t);
            return h.found;
        }
        std::vector<Hit> hits;
        AllHits<R>(s, o, d, R(0.001), tmax, hits);
        for (const Hit& h : hits) {
            if (Opacity(h.primitive, R(h.u), R(h.v)) >= R(0.5)) {
                hitDistance = R(h.t);
                return true;
            }
        }
        return false;
    }
    // shadowVisibility (:855-889) -> RGB visibility.
    V3<R> ShadowVisibility(const V3<R>& o, const V3<R>& d, R maxDist, const Light& L) const
    {
        if (L.shadow.x < 0.5F) return V3<R>(R(1));
        R dist = maxDist;
        if (L.shadow.y >= 0.0F) dist = std::min(dist, R(L.shadow.y));
        if (dist <= R(0.001)) return V3<R>(R(1));
        R hitDistance = R(0);
        if (!Occluded(o, d, dist, hitDistance)) return V3<R>(R(1));
        R strength = R(1);
        if (L.shadow.y >= 0.0F && L.shadow.z > 0.0F) {
            const R falloff = std::min(R(L.shadow.z), std::max(R(L.shadow.y), R(0)));
            strength = clampR((R(L.shadow.y) - hitDistance) / std::max(falloff, R(1e-6)), R(0), R(1));
            strength = std::pow(strength, std::max(R(L.shadow.w), R(0)));
        }
        const V3<R> sc = XYZ(L.shadowColor);
        return V3<R>(mixR(R(1), sc.x, strength), mixR(R(1), sc.y, strength), mixR(R(1), sc.z, strength));
    }
    // meshLightVisible (:891-912)
    bool MeshLightVisible(const V3<R>& o, const V3<R>& d, R maxDist) const
    {
        if (maxDist <= R(0.001)) return true;
        R hd;
        return !Occluded(o, d, maxDist, hd);
    }
    // sampleEmissiveTriangle (:914-924)
    std::uint32_t SampleEmissiveTriangle(R target) const
    {
        std::uint32_t lo = 0, hi = s.triangleCount - 1U;
        while (lo < hi) {
            const std::uint32_t mid = lo + (hi - lo) / 2U;
            if (target <= R(s.emissive[mid * 4])) hi = mid; else lo = mid + 1U;
        }
        return lo;
    }
    // emissiveTrianglePdf (:926-931)
    R EmissivePdf(std::uint32_t prim, R d2, R cosine) const
    {
        if (cosine <= R(0)) return R(0);
        return R(s.emissive[prim * 4 + 1]) * d2 / cosine;
    }
    // eonDirectionalAlbedo (:933-944)
    static R EonAlbedo(R nd, R rough)
    {
        const R c1 = R(0.2877934092);
        const R c = clampR(nd, R(0), R(1));
        const R sn = std::sqrt(std::max(R(1) - c * c, R(0)));
        const R theta = std::acos(c);
        const R s3 = sn * sn * sn;
        const R g = sn * (theta - sn * c) + (R(2) / R(3)) * (sn * (R(1) - s3) / std::max(c, R(1e-5)) - sn);
        const R a = R(1) / (R(1) + c1 * rough);
        return clampR(a + rough * a * g / Pi(), R(0), R(1));
    }
    // evaluateDiffuseReflection (:946-1009)
    static R EvalDiffuse(const V3<R>& n, const V3<R>& v, const V3<R>& l, R color, R rough, std::uint32_t model)
    {
        const R nl = std::max(dot(n, l), R(0)), nv = std::max(dot(n, v), R(0));
        if (nl <= R(0) || nv <= R(0)) return R(0);
        if (model == 1U) {
            const R sg = clampR(rough, R(0), R(1));
            const R s2 = sg * sg;
            const R a = R(1) - R(0.5) * s2 / (s2 + R(0.33));
            const R b = R(0.45) * s2 / (s2 + R(0.09));
            const V3<R> pl = l - n * nl, pv = v - n * nv;
            R az = R(0);
            if (dot(pl, pl) > R(1e-8) && dot(pv, pv) > R(1e-8)) az = std::max(dot(normalize(pl), normalize(pv)), R(0));
            const R sl = std::sqrt(std::max(R(1) - nl * nl, R(0))), sv = std::sqrt(std::max(R(1) - nv * nv, R(0)));
            const R sa = std::max(sl, sv);
            const R tb = std::min(sl / std::max(nl, R(1e-5)), sv / std::max(nv, R(1e-5)));
            return nl * (a + b * az * sa * tb) / Pi();
        }
        if (model == 2U) {
            const V3<R> h = normalize(v + l);
            const R lh = std::max(dot(l, h), R(0));
            const R grazing = R(0.5) + R(2) * clampR(rough, R(0), R(1)) * lh * lh;
            const R ls = R(1) + (grazing - R(1)) * std::pow(R(1) - nl, R(5));
            const R vs = R(1) + (grazing - R(1)) * std::pow(R(1) - nv, R(5));
            return nl * ls * vs / Pi();
        }
        if (model == 3U) {
            const R c1 = R(0.2877934092), c2 = R(0.0726760455);
            const R sg = clampR(rough, R(0), R(1));
            const R a = R(1) / (R(1) + c1 * sg);
            const R sv = dot(l, v) - nl * nv;
            const R st = sv > R(0) ? sv / std::max(nl, nv) : sv;
            const R single = a * (R(1) + sg * st);
            const R avg = (R(1) + c2 * sg) / (R(1) + c1 * sg);
            const R eo = EonAlbedo(nv, sg), ei = EonAlbedo(nl, sg);
            const R dirf = (R(1) - eo) * (R(1) - ei) / std::max(R(1) - avg, R(1e-5));
            const R rho = clampR(color, R(0), R(0.9999));
            const R ms = rho * avg / std::max(R(1) - rho * (R(1) - avg), R(1e-5));
            return nl / Pi() * (single + ms * dirf);
        }
        return nl / Pi();
    }

    struct Closure {
        R diffuseColor, transmission, subsurface, diffuseWeight, diffuseRoughness;
        std::uint32_t diffuseModel;
        R translucentWeight, translucentColor, metalness, rx, ry, f0, coat, crx, cry, coatF0, sheen, sheenColor,
            sheenRoughness;
        bool sheenZeltner;
    };
    // evaluateDirectSurface (:1013-1069)
    static R EvalDirect(const V3<R>& n, const V3<R>& t, const V3<R>& v, const V3<R>& l, const Closure& c,
                        R diffuseScale, R specularScale)
    {
        const R snl = dot(n, l);
        const R nl = std::max(snl, R(0));
        const R bnl = std::max(-snl, R(0));
        const R wrapped = std::max((snl + R(0.5) * c.subsurface) / (R(1) + R(0.5) * c.subsurface), R(0));
        if (wrapped <= R(0) && nl <= R(0) && bnl * c.translucentWeight <= R(0)) return R(0);
        const V3<R> h = nl > R(0) ? normalize(v + l) : n;
        const R baseF = Fresnel(std::max(dot(v, h), R(0)), c.f0);
        const R dresp = mixR(EvalDiffuse(n, v, l, c.diffuseColor, c.diffuseRoughness, c.diffuseModel),
                             wrapped / Pi(), c.subsurface);
        const R diffuse = diffuseScale * (R(1) - baseF) * c.diffuseWeight * (R(1) - c.translucentWeight) *
                          (R(1) - c.metalness) * c.diffuseColor * dresp * (R(1) - c.transmission);
        const R viewF = Fresnel(std::max(dot(n, v), R(0)), c.f0);
        const R translucent = diffuseScale * (R(1) - viewF) * c.translucentWeight * (R(1) - c.metalness) *
                              c.translucentColor * (bnl / Pi()) * (R(1) - c.transmission);
        const R specular = specularScale * nl * EvalGGX(n, t, v, l, c.rx, c.ry, c.f0);
        R coatF = R(0), coatSpec = R(0);
        if (c.coat > R(0)) {
            coatF = Fresnel(std::max(dot(v, h), R(0)), c.coatF0);
            coatSpec = specularScale * c.coat * nl * EvalGGX(n, t, v, l, c.crx, c.cry, c.coatF0);
        }
        const R baseLayer = (R(1) - c.coat * coatF) * (diffuse + translucent + specular) + coatSpec;
        const R nv = std::max(dot(n, v), R(0));
        const R sheenAlbedo = c.sheenZeltner ? SheenAlbedoZ(nv, c.sheenRoughness) : SheenAlbedoCK(nv, c.sheenRoughness);
        const R sheenLayer = specularScale * c.sheen * c.sheenColor * SheenResponse(n, v, l, c.sheenRoughness, c.sheenZeltner);
        return baseLayer * (R(1) - clampR(c.sheen * sheenAlbedo, R(0), R(1))) + sheenLayer;
    }
    // offsetSurfaceOrigin (:1071-1075)
    static V3<R> Offset(const V3<R>& p, const V3<R>& gn, const V3<R>& d)
    {
        const R side = dot(gn, d) >= R(0) ? R(1) : R(-1);
        return p + gn * (R(0.002) * side);
    }
    // shapingConeFactor (:1085-1094)
    static R ShapingCone(const Light& L, R cosine)
    {
        const R angle = std::acos(clampR(cosine, R(-1), R(1)));
        const R outer = clampR(R(L.shaping.y), R(0), R(180)) * (Pi() / R(180));
        const R inner = outer * (R(1) - clampR(R(L.shaping.z), R(0), R(1)));
        if (outer >= R(3.1415925)) return R(1);
        if (outer <= inner + R(1e-6)) return angle <= outer ? R(1) : R(0);
        return R(1) - smoothstepR(inner, outer, angle);
    }
    // dispersedIor (:1096-1110)
    static R DispersedIor(R iorD, R scale, R abbe, R w)
    {
        if (scale <= R(0) || abbe <= R(1)) return iorD;
        const R lf = R(486.13), ld = R(587.56), lc = R(656.27);
        const R rd = R(1) / (lf * lf) - R(1) / (lc * lc);
        const R b = (iorD - R(1)) / std::max(abbe * rd, R(1e-8));
        const R a = iorD - b / (ld * ld);
        const R cauchy = a + b / (w * w);
        return mixR(iorD, cauchy, clampR(scale, R(0), R(1)));
    }
    // sampleHenyeyGreenstein (:1112-1128)
    static V3<R> SampleHG(const V3<R>& fwd, R aniso, std::uint32_t& st)
    {
        const R g = clampR(aniso, R(-0.95), R(0.95));
        const R u = Rand(st);
        const R cosine = std::fabs(g) < R(1e-3)
            ? R(1) - R(2) * u
            : (R(1) + g * g - sq((R(1) - g * g) / (R(1) - g + R(2) * g * u))) / (R(2) * g);
        const R sine = std::sqrt(std::max(R(1) - cosine * cosine, R(0)));
        const R phi = TwoPi() * Rand(st);
        const V3<R> w = normalize(fwd);
        const V3<R> t = AnyTangent(w);
        const V3<R> b = cross(w, t);
        return normalize(t * (std::cos(phi) * sine) + b * (std::sin(phi) * sine) + w * cosine);
    }
    // Reflection or refraction at a dielectric boundary (:1737-1750): Schlick
    // Fresnel from the spectral IOR picks the event with one random number
    // (after TIR, which always reflects); thin walls pass straight through.
    // rd becomes the new direction; returns true when the path crosses into
    // or out of the medium.
    static bool Transmit(V3<R>& rd, const V3<R>& normal, bool front, R ior, bool thin, std::uint32_t& rng)
    {
        const R eta = front ? R(1) / ior : ior;
        const R cosT = std::min(dot(-rd, normal), R(1));
        R r0 = (R(1) - ior) / (R(1) + ior);
        r0 *= r0;
        const R fres = r0 + (R(1) - r0) * std::pow(R(1) - cosT, R(5));
        const V3<R> refr = refract(rd, normal, eta);
        const bool tir = dot(refr, refr) < R(1e-8);
        if (tir || Rand(rng) < fres) {
            rd = reflect(rd, normal);
            return false;
        }
        rd = thin ? rd : normalize(refr);
        return !thin;
    }
    // traceOpaqueBoundary (:1130-1151)
    bool OpaqueBoundary(const V3<R>& o, const V3<R>& d, V3<R>& boundary, V3<R>& bn, R& dist) const
    {
        const Hit h = ClosestHit<R>(s, o, d, R(0.001), R(10000));
        if (!h.found) return false;
        const std::uint32_t p = h.primitive;
        dist = R(h.t);
        boundary = o + d * dist;
        const V3<R> p0 = s.Pos<R>(s.indices[p * 3]), p1 = s.Pos<R>(s.indices[p * 3 + 1]),
                    p2 = s.Pos<R>(s.indices[p * 3 + 2]);
        bn = normalize(cross(p1 - p0, p2 - p0));
        if (dot(bn, d) > R(0)) bn = -bn;
        return true;
    }
    // randomWalkSubsurface (:1153-1186)
    bool RandomWalk(const V3<R>& entry, const V3<R>& outN, R mfp, R albedo, R aniso, std::uint32_t& st,
                    V3<R>& exitO, V3<R>& exitD, R& atten) const
    {
        const R radius = std::max(mfp, R(0.001));
        const R sigmaS = R(1) / radius;
        const R sigmaA = -std::log(clampR(albedo, R(0.001), R(0.9999))) / radius;
        V3<R> o = entry - outN * R(0.003);
        V3<R> d = CosineHemisphere(-outN, st);
        atten = R(1);
        for (int step = 0; step < 8; ++step) {
            V3<R> boundary, bn;
            R bd;
            if (!OpaqueBoundary(o, d, boundary, bn, bd)) return false;
            const R scatter = -std::log(std::max(R(1) - Rand(st), R(1e-6))) / sigmaS;
            const bool force = step == 7;
            const R travel = force ? bd : std::min(scatter, bd);
            atten *= std::exp(-sigmaA * travel);
            if (force || scatter >= bd) {
                exitD = d;
                exitO = boundary + d * R(0.003);
                return true;
            }
            o += d * scatter;
            d = SampleHG(d, aniso, st);
        }
        return false;
    }

    // Primary closest hit of the path loop (:1228-1243): opaque scenes take
    // the closest triangle; otherwise candidates in (t, primitive) order are
    // confirmed with randomFloat <= opacity.
    Hit PathHit(const V3<R>& o, const V3<R>& d, std::uint32_t& rng) const
    {
        if (opaqueFlag) return ClosestHit<R>(s, o, d, R(0.001), R(10000));
        std::vector<Hit> hits;
        AllHits<R>(s, o, d, R(0.001), R(10000), hits);
        for (const Hit& h : hits)
            if (Rand(rng) <= Opacity(h.primitive, R(h.u), R(h.v))) return h;
        return Hit{};
    }

    // One lane path of `main` (:1198-1881). Returns the lane's scalar
    // radiance (radiance.r) and its wavelength.
    R TraceLane(const hdcx_camera& cam, std::uint32_t width, std::uint32_t height, std::uint32_t px,
                std::uint32_t py, std::uint32_t sampleNumber, std::uint32_t lane, int maxBounces, float& wlOut) const
    {
        const std::uint32_t index = py * width + px;
        std::uint32_t rng = HashState(index ^ (sampleNumber * 0x9e3779b9U) ^ 0xa511e9b3U);
        const float shift = static_cast<float>(HashState(index ^ 0x68bc21ebU)) * (1.0F / 4294967296.0F);
        const float lanePhase = (static_cast<float>(sampleNumber & 31U) + 0.5F) / 32.0F;
        const float laneStep = static_cast<float>(lane) / 3.0F;
        const float fr = shift + lanePhase + laneStep;
        const float wlf = 380.0F + 400.0F * (fr - std::floor(fr));
        wlOut = wlf;
        const R wl = R(wlf);
        const R jx = Rand(rng), jy = Rand(rng);
        const R u = (R(px) + jx) / R(width), v = (R(py) + jy) / R(height);
        const V3<R> org(R(cam.origin[0]), R(cam.origin[1]), R(cam.origin[2]));
        const V3<R> hz(R(cam.horizontal[0]), R(cam.horizontal[1]), R(cam.horizontal[2]));
        const V3<R> vt(R(cam.vertical[0]), R(cam.vertical[1]), R(cam.vertical[2]));
        V3<R> ro = org;
        // (:1211-1213) with lowerLeft - origin evaluated once in FP64 and
        // rounded to R: the same direction without the FP32 cancellation
        // that costs ~7 bits when |origin| is large against the image plane
        // distance (the FP64 oracle is unchanged by this).
        const V3<R> toLL(R(double(cam.lower_left[0]) - double(cam.origin[0])),
                         R(double(cam.lower_left[1]) - double(cam.origin[1])),
                         R(double(cam.lower_left[2]) - double(cam.origin[2])));
        V3<R> rd = normalize(toLL + hz * u + vt * v);
        R throughput = R(1), radiance = R(0);
        bool envOnMiss = true, prevDelta = true;
        R prevBsdfPdf = R(0), prevEnvPdf = R(0);
        bool inMedium = false;
        R mAbs = R(0), mScat = R(0), mAniso = R(0);
        const bool hasMeshLights = HasEmissive();

        for (int bounce = 0; bounce < maxBounces; ++bounce) {
            const Hit hit = PathHit(ro, rd, rng);
            if (!hit.found) {
                if (envOnMiss) {
                    const R mis = prevDelta ? R(1) : Power(prevBsdfPdf, prevEnvPdf);
                    radiance += throughput * mis * Illuminant(Environment(rd), wl);
                }
                break;
            }
            const std::uint32_t prim = hit.primitive;
            const V3<R> p0 = s.Pos<R>(s.indices[prim * 3]), p1 = s.Pos<R>(s.indices[prim * 3 + 1]),
                        p2 = s.Pos<R>(s.indices[prim * 3 + 2]);
            const V3<R> gn = normalize(cross(p1 - p0, p2 - p0));
            const bool front = dot(gn, rd) < R(0);
            const V3<R> ogn = front ? gn : -gn;
            const R bu = R(hit.u), bv = R(hit.v);
            const V3<R> authored = ShadingNormal(prim, bu, bv, gn);
            V3<R> normal = front ? authored : -authored;
            const float* tc = &s.texcoords[prim * 6];
            const R u0 = tc[0], v0 = tc[1], u1 = tc[2], v1 = tc[3], u2 = tc[4], v2 = tc[5];
            R su, sv;
            SurfaceUv(prim, bu, bv, su, sv);
            const R distance = R(hit.t);
            const V3<R> hp = ro + rd * distance;
            if (inMedium) {
                if (mScat > R(0)) {
                    const R sd = -std::log(std::max(R(1) - Rand(rng), R(1e-6))) / mScat;
                    if (sd < distance) {
                        throughput *= std::exp(-mAbs * sd);
                        ro += rd * sd;
                        rd = SampleHG(rd, mAniso, rng);
                        envOnMiss = true;
                        prevDelta = true;
                        prevBsdfPdf = R(0);
                        continue;
                    }
                }
                throughput *= std::exp(-mAbs * distance);
            }
            const Material& m = MaterialOf(prim);
            // Material fetch (:1296-1437).
            V3<R> baseRgb = XYZ(m.baseColorMetalness);
            if (m.textureIndices0.x != kMissing) {
                const R fb[4] = {baseRgb.x, baseRgb.y, baseRgb.z, R(1)};
                R smp[4];
                SampleMat(m.textureIndices0.x, su, sv, fb, smp);
                const int channel = static_cast<int>(std::round(m.baseColorTextureScale.w));
                const V3<R> src = channel < 0 ? V3<R>(smp[0], smp[1], smp[2]) : V3<R>(smp[std::min(channel, 3)]);
                baseRgb = XYZ(m.baseColorTextureBias) + XYZ(m.baseColorTextureScale) * src;
            }
            const R baseColor = Reflectance(baseRgb, wl);
            const R metalness = clampR(SampleMatR(m.textureIndices0.y, su, sv, R(m.baseColorMetalness.w)), R(0), R(1));
            R rx, ry;
            {
                const R fb[4] = {R(m.emissionRoughness.w), R(m.subsurfaceColor.w), R(0), R(1)};
                R o[4];
                SampleMat(m.textureIndices0.z, su, sv, fb, o);
                rx = clampR(o[0], R(0.02), R(1));
                ry = clampR(o[1], R(0.02), R(1));
            }
            const R diffuseRoughness = clampR(SampleMatR(m.textureIndices4.z, su, sv, R(m.diffuseProperties.x)), R(0), R(1));
            const R diffuseWeight = clampR(SampleMatR(m.textureIndices4.w, su, sv, R(m.diffuseProperties.z)), R(0), R(1));
            const std::uint32_t diffuseModel = static_cast<std::uint32_t>(m.diffuseProperties.y + 0.5F);
            normal = NormalMap(m.textureIndices1.y, m.coatWeightRoughnessIor.w > 0.5F, R(m.subsurfaceWeightScale.w), su, sv,
                               normal, p0, p1, p2, u0, v0, u1, v1, u2, v2);
            const V3<R> tangent = SurfaceTangent(normal, p0, p1, p2, u0, v0, u1, v1, u2, v2);
            const V3<R> emissionRgb = EmissionRgb(prim, su, sv);
            const R emission = Illuminant(emissionRgb, wl);
            R emissionMis = R(1);
            if (!prevDelta && (emissionRgb.x > R(0) || emissionRgb.y > R(0) || emissionRgb.z > R(0))) {
                const R ec = std::fabs(dot(gn, -rd));
                const R lp = EmissivePdf(prim, distance * distance, ec);
                emissionMis = Power(prevBsdfPdf, lp);
            }
            radiance += throughput * emission * emissionMis;

            const R transmission = clampR(R(m.transmissionOpacityIor.x), R(0), R(1));
            V3<R> transColorRgb = XYZ(m.transmissionColorThinWalled);
            if (transmission > R(0)) transColorRgb = SampleMatRgb(m.textureIndices1.z, su, sv, transColorRgb);
            const R transmissionColor = Reflectance(transColorRgb, wl);
            R subsurface = R(0);
            R diffuseColor = baseColor;
            V3<R> sssColorRgb = XYZ(m.subsurfaceColor);
            V3<R> sssRadiusRgb = vmax(XYZ(m.subsurfaceRadius), R(0.001));
            if (m.subsurfaceWeightScale.x > 0.0F || m.textureIndices2.x != kMissing) {
                subsurface = clampR(SampleMatR(m.textureIndices2.x, su, sv, R(m.subsurfaceWeightScale.x)), R(0), R(1));
                if (subsurface > R(0)) {
                    sssColorRgb = SampleMatRgb(m.textureIndices2.y, su, sv, XYZ(m.subsurfaceColor));
                    sssRadiusRgb = vmax(SampleMatRgb(m.textureIndices2.z, su, sv, XYZ(m.subsurfaceRadius)), R(0.001));
                    const R sssColor = Reflectance(sssColorRgb, wl);
                    diffuseColor = mixR(baseColor, sssColor, subsurface);
                }
            }
            const R ior = DispersedIor(std::max(R(m.transmissionOpacityIor.z), R(1.0001)), R(m.transmissionMedium.z),
                                       R(m.transmissionMedium.w), wl);
            const R specularWeight = clampR(SampleMatR(m.textureIndices3.x, su, sv, R(m.specularWeightColor.x)), R(0), R(1));
            const V3<R> specColorRgb = SampleMatRgb(m.textureIndices3.y, su, sv,
                                                    V3<R>(R(m.specularWeightColor.y), R(m.specularWeightColor.z),
                                                          R(m.specularWeightColor.w)));
            const R specularColor = Reflectance(specColorRgb, wl);
            R dielectric = (ior - R(1)) / (ior + R(1));
            dielectric *= dielectric;
            const R f0 = mixR(clampR(dielectric * specularWeight * specularColor, R(0), R(1)), baseColor, metalness);
            R coat = R(0);
            R crx = R(m.coatWeightRoughnessIor.y), cry = R(m.subsurfaceRadius.w);
            R coatF0 = R(0);
            if (m.coatWeightRoughnessIor.x > 0.0F || m.textureIndices3.z != kMissing) {
                coat = clampR(SampleMatR(m.textureIndices3.z, su, sv, R(m.coatWeightRoughnessIor.x)), R(0), R(1));
                if (coat > R(0)) {
                    const V3<R> ccRgb = SampleMatRgb(m.textureIndices3.w, su, sv, XYZ(m.coatColor));
                    const R coatColor = Reflectance(ccRgb, wl);
                    const R fb[4] = {crx, cry, R(0), R(1)};
                    R o[4];
                    SampleMat(m.textureIndices4.x, su, sv, fb, o);
                    crx = clampR(o[0], R(0.02), R(1));
                    cry = clampR(o[1], R(0.02), R(1));
                    const R cior = std::max(R(m.coatWeightRoughnessIor.z), R(1.0001));
                    R cr = (cior - R(1)) / (cior + R(1));
                    cr *= cr;
                    coatF0 = clampR(cr * coatColor, R(0), R(1));
                }
            }
            const R sheen = clampR(SampleMatR(m.textureIndices1.w, su, sv, R(m.coatColor.w)), R(0), R(1));
            const V3<R> sheenRgb = SampleMatRgb(m.textureIndices2.w, su, sv, XYZ(m.sheenColorMode));
            const R sheenColor = Reflectance(sheenRgb, wl);
            const R sheenRoughness = clampR(SampleMatR(m.textureIndices4.y, su, sv, R(m.transmissionScatter.w)), R(0.01), R(1));
            const bool sheenZ = m.sheenColorMode.w > 0.5F;
            const R translucentWeight = clampR(SampleMatR(m.textureIndices5.x, su, sv, R(m.translucentColorWeight.w)), R(0), R(1));
            const V3<R> tlRgb = SampleMatRgb(m.textureIndices5.y, su, sv, XYZ(m.translucentColorWeight));
            const R translucentColor = Reflectance(tlRgb, wl);

            const V3<R> view = normalize(-rd);
            const R nv = std::max(dot(normal, view), R(0));
            const R sheenAlbedo = sheenZ ? SheenAlbedoZ(nv, sheenRoughness) : SheenAlbedoCK(nv, sheenRoughness);
            const R pSheen = clampR(sheen * sheenAlbedo * LumGrey(sheenColor), R(0), R(0.95));
            const R coatContribution = coat > R(0) ? coat * Fresnel(nv, coatF0) : R(0);
            const R pCoat = clampR(LumGrey(coatContribution), R(0), R(0.95));
            const R baseF = Fresnel(nv, f0);
            const R specContribution = baseF;
            const R diffContribution = (R(1) - baseF) * diffuseWeight * (R(1) - translucentWeight) * (R(1) - metalness) * diffuseColor;
            const R tlContribution = (R(1) - baseF) * translucentWeight * (R(1) - metalness) * translucentColor;
            const R specE = LumGrey(specContribution);
            const R diffE = LumGrey(diffContribution);
            const R tlE = LumGrey(tlContribution);
            const R baseE = diffE + tlE;
            const R pSpec = clampR(specE / std::max(specE + baseE, R(1e-5)), R(0.05), R(0.95));
            const R pTl = baseE > R(1e-5) ? tlE / baseE : R(0);
            const Closure cl{diffuseColor, transmission, subsurface, diffuseWeight, diffuseRoughness, diffuseModel,
                             translucentWeight, translucentColor, metalness, rx, ry, f0, coat, crx, cry, coatF0,
                             sheen, sheenColor, sheenRoughness, sheenZ};

            // Analytic-light NEE (:1469-1592).
            if (lightCount > 0U) {
                const R dSpecE = specE + LumGrey(coatContribution) + pSheen;
                std::uint32_t li;
                R selPdf;
                SampleAnalyticLight(hp, normal, diffE, dSpecE, tlE, rng, li, selPdf);
                const Light& L = s.lights[li];
                V3<R> ld(0, 1, 0);
                V3<R> lrad(R(0));
                R invPdf = R(0);
                R maxShadow = R(10000);
                if (L.textureInfo.z == 0U) {
                    R lp = R(0);
                    SampleDomeLight(L, rng, ld, lp);
                    lrad = SampleDome(L, ld);
                    invPdf = lp > R(0) ? R(1) / lp : R(0);
                } else if (L.textureInfo.z == 5U) {
                    const R half = R(0.5) * (clampR(R(L.geometry.z), R(0), R(359.999)) * (Pi() / R(180)));
                    if (half <= R(1e-6)) {
                        ld = normalize(XYZ(L.basisZ));
                        invPdf = R(1);
                    } else {
                        const R cmax = std::cos(half);
                        const R ct = mixR(cmax, R(1), Rand(rng));
                        const R stt = std::sqrt(std::max(R(1) - ct * ct, R(0)));
                        const R phi = TwoPi() * Rand(rng);
                        ld = normalize(XYZ(L.basisX) * (std::cos(phi) * stt) + XYZ(L.basisY) * (std::sin(phi) * stt) +
                                       XYZ(L.basisZ) * ct);
                        invPdf = R(1);
                    }
                    lrad = XYZ(L.colorIntensity) * R(L.colorIntensity.w);
                } else {
                    const R lu = Rand(rng), lv = Rand(rng);
                    V3<R> lpos = XYZ(L.position);
                    V3<R> ln = -normalize(XYZ(L.basisZ));
                    if (L.textureInfo.z == 1U) {
                        lpos += XYZ(L.axisU) * (lu - R(0.5)) + XYZ(L.axisV) * (lv - R(0.5));
                    } else if (L.textureInfo.z == 2U) {
                        const R dr = std::sqrt(lu), dphi = TwoPi() * lv;
                        lpos += (XYZ(L.axisU) * std::cos(dphi) + XYZ(L.axisV) * std::sin(dphi)) * dr;
                    } else if (L.textureInfo.z == 3U) {
                        ln = UniformSphere(rng);
                        lpos += ln * R(L.geometry.x);
                    } else if (L.textureInfo.z == 4U) {
                        const R cphi = TwoPi() * lv;
                        ln = normalize(XYZ(L.basisY) * std::cos(cphi) + XYZ(L.basisZ) * std::sin(cphi));
                        lpos += XYZ(L.axisU) * (lu - R(0.5)) + ln * R(L.geometry.x);
                    }
                    const V3<R> toL = lpos - hp;
                    const R d2 = dot(toL, toL);
                    const R dist = std::sqrt(std::max(d2, R(1e-8)));
                    ld = toL / dist;
                    const R ec = std::max(dot(ln, -ld), R(0));
                    if (ec > R(0) && L.controls.w > 0.0F) {
                        const R shapingCos = std::max(dot(normalize(XYZ(L.basisZ)), ld), R(0));
                        const R focus = std::pow(shapingCos, std::max(R(L.shaping.x), R(0)));
                        const V3<R> tint = XYZ(L.focusTint);
                        const V3<R> focusColor(mixR(tint.x, R(1), focus), mixR(tint.y, R(1), focus), mixR(tint.z, R(1), focus));
                        V3<R> texv(R(1));
                        if (L.textureInfo.z == 1U && L.textureInfo.x != kMissing) texv = TexRgb(L.textureInfo.x, lu, lv);
                        lrad = XYZ(L.colorIntensity) * R(L.colorIntensity.w) * texv * focusColor * ShapingCone(L, shapingCos);
                        invPdf = R(L.controls.w) * ec / std::max(d2, R(1e-8));
                        maxShadow = std::max(dist - R(0.003), R(0.001));
                    }
                }
                if (invPdf > R(0)) {
                    const R direct = EvalDirect(normal, tangent, view, ld, cl, R(L.controls.x), R(L.controls.y));
                    if (direct > R(0)) {
                        const V3<R> vis = ShadowVisibility(Offset(hp, ogn, ld), ld, maxShadow, L);
                        const R sLight = Illuminant(lrad, wl);
                        const R sVis = Reflectance(vis, wl);
                        const R mis = L.textureInfo.z == 0U
                            ? Power(EnvLightPdf(hp, normal, ld, diffE, dSpecE, tlE),
                                    SurfacePdf(normal, tangent, view, ld, rx, ry, crx, cry, pSheen, pCoat, pSpec, pTl))
                            : R(1);
                        radiance += throughput * direct * sLight * sVis * invPdf * mis / std::max(selPdf, R(1e-8));
                    }
                }
            } else if (!hasMeshLights) {
                // Fallback sun and sky (:1593-1668).
                const V3<R> up = zUp ? V3<R>(0, 0, 1) : V3<R>(0, 1, 0);
                const V3<R> az = zUp ? normalize(V3<R>(1, 1, 0)) : normalize(V3<R>(1, 0, 1));
                const V3<R> sunDir = normalize(up * R(0.9659258) + az * R(0.2588190));
                const R direct = EvalDirect(normal, tangent, view, sunDir, cl, R(1), R(1));
                if (direct > R(0)) {
                    const V3<R> vis = ShadowVisibility(Offset(hp, ogn, sunDir), sunDir, R(10000), s.lights[0]);
                    const R sSun = Illuminant(V3<R>(R(3.45575), R(3.14159), R(2.82743)), wl);
                    radiance += throughput * direct * sSun * Reflectance(vis, wl);
                }
                const V3<R> skyDir = CosineHemisphere(normal, rng);
                const R skyDirect = EvalDirect(normal, tangent, view, skyDir, cl, R(1), R(0));
                if (skyDirect > R(0)) {
                    const V3<R> vis = ShadowVisibility(Offset(hp, ogn, skyDir), skyDir, R(10000), s.lights[0]);
                    const R sSky = Illuminant(Environment(skyDir), wl);
                    radiance += throughput * skyDirect * sSky * Reflectance(vis, wl) *
                                (Pi() / std::max(dot(normal, skyDir), R(1e-4)));
                }
                if (translucentWeight > R(0)) {
                    const V3<R> backDir = CosineHemisphere(-normal, rng);
                    const R backDirect = EvalDirect(normal, tangent, view, backDir, cl, R(1), R(0));
                    if (backDirect > R(0)) {
                        const V3<R> vis = ShadowVisibility(Offset(hp, ogn, backDir), backDir, R(10000), s.lights[0]);
                        const R sSky = Illuminant(Environment(backDir), wl);
                        radiance += throughput * backDirect * sSky * Reflectance(vis, wl) *
                                    (Pi() / std::max(dot(-normal, backDir), R(1e-4)));
                    }
                }
            }
            // Emissive-mesh NEE (:1670-1731).
            if (hasMeshLights) {
                const std::uint32_t lp = SampleEmissiveTriangle(Rand(rng));
                if (lp != prim && s.emissive[lp * 4 + 1] > 0.0F) {
                    const R root = std::sqrt(Rand(rng));
                    const R second = Rand(rng);
                    const R b0 = R(1) - root, b1 = root * (R(1) - second), b2 = root * second;
                    const V3<R> q0 = s.Pos<R>(s.indices[lp * 3]), q1 = s.Pos<R>(s.indices[lp * 3 + 1]),
                                q2 = s.Pos<R>(s.indices[lp * 3 + 2]);
                    const V3<R> lpos = q0 * b0 + q1 * b1 + q2 * b2;
                    const V3<R> ln = normalize(cross(q1 - q0, q2 - q0));
                    const V3<R> toL = lpos - hp;
                    const R d2 = dot(toL, toL);
                    const R dist = std::sqrt(std::max(d2, R(1e-8)));
                    const V3<R> ld = toL / dist;
                    const R ec = std::fabs(dot(ln, -ld));
                    const R lightPdf = EmissivePdf(lp, d2, ec);
                    if (lightPdf > R(0)) {
                        const R direct = EvalDirect(normal, tangent, view, ld, cl, R(1), R(1));
                        if (direct > R(0) && MeshLightVisible(Offset(hp, ogn, ld), ld, std::max(dist - R(0.003), R(0.001)))) {
                            R lu, lv;
                            SurfaceUv(lp, b1, b2, lu, lv);
                            V3<R> le = EmissionRgb(lp, lu, lv);
                            le = le * Opacity(lp, b1, b2);
                            const R sLight = Illuminant(le, wl);
                            const R bsdfPdf = SurfacePdf(normal, tangent, view, ld, rx, ry, crx, cry, pSheen, pCoat, pSpec, pTl);
                            const R mis = Power(lightPdf, bsdfPdf);
                            radiance += throughput * direct * sLight * (mis / lightPdf);
                        }
                    }
                }
            }
            // Transmission (:1735-1777).
            if (Rand(rng) < transmission) {
                const bool through = Transmit(rd, normal, front, ior, m.transmissionColorThinWalled.w > 0.5F, rng);
                const R depth = std::max(R(m.transmissionMedium.x), R(0));
                throughput *= depth > R(0) ? R(1) : transmissionColor;
                if (through) {
                    if (front) {
                        const R st = std::max(transmissionColor, R(0.001));
                        mAbs = depth > R(0) ? -std::log(st) / depth : R(0);
                        const R scat = std::max(ReconstructRgb(XYZ(m.transmissionScatter), wl), R(0));
                        mScat = depth > R(0) ? scat / depth : scat;
                        mAniso = R(m.transmissionMedium.y);
                        inMedium = true;
                    } else {
                        inMedium = false;
                        mAbs = R(0);
                        mScat = R(0);
                    }
                }
                ro = hp + rd * R(0.002);
                envOnMiss = true;
                prevDelta = true;
                prevBsdfPdf = R(0);
                continue;
            }
            // BSDF sampling chain (:1779-1858).
            envOnMiss = false;
            ro = hp + ogn * R(0.002);
            if (pSheen > R(0) && Rand(rng) < pSheen) {
                rd = CosineHemisphere(normal, rng);
                const R sp = std::max(dot(normal, rd), R(0)) / Pi();
                const R resp = SheenResponse(normal, view, rd, sheenRoughness, sheenZ);
                throughput *= sheen * sheenColor * resp / std::max(pSheen * sp, R(1e-6));
                envOnMiss = true;
            } else {
                throughput *= (R(1) - sheen * sheenAlbedo) / std::max(R(1) - pSheen, R(1e-4));
                if (pCoat > R(0) && Rand(rng) < pCoat) {
                    const V3<R> inc = rd;
                    rd = SampleGGX(rd, normal, tangent, crx, cry, rng);
                    throughput *= coat * GGXWeight(inc, rd, normal, tangent, crx, cry, coatF0) / pCoat;
                    envOnMiss = std::max(crx, cry) < R(0.2);
                } else {
                    throughput *= (R(1) - coatContribution) / std::max(R(1) - pCoat, R(1e-4));
                    if (Rand(rng) < pSpec) {
                        const V3<R> inc = rd;
                        rd = SampleGGX(rd, normal, tangent, rx, ry, rng);
                        throughput *= GGXWeight(inc, rd, normal, tangent, rx, ry, f0) / pSpec;
                        envOnMiss = std::max(rx, ry) < R(0.2);
                    } else {
                        if (pTl > R(0) && Rand(rng) < pTl) {
                            throughput *= tlContribution / std::max((R(1) - pSpec) * pTl, R(1e-4));
                            rd = CosineHemisphere(-normal, rng);
                            ro = Offset(hp, ogn, rd);
                        } else {
                            throughput *= diffContribution / std::max((R(1) - pSpec) * (R(1) - pTl), R(1e-4));
                            if (subsurface > R(0) && Rand(rng) < subsurface) {
                                const R radius = std::max(ReconstructRgb(sssRadiusRgb, wl) *
                                                              std::max(R(m.subsurfaceWeightScale.y), R(0)),
                                                          R(0.001));
                                const R albedo = Reflectance(sssColorRgb, wl);
                                V3<R> eo, ed;
                                R att;
                                if (RandomWalk(hp, ogn, radius, albedo, R(m.subsurfaceWeightScale.z), rng, eo, ed, att)) {
                                    throughput *= att;
                                    ro = eo;
                                    rd = ed;
                                    envOnMiss = true;
                                    prevDelta = true;
                                    prevBsdfPdf = R(0);
                                    continue;
                                }
                            }
                            rd = CosineHemisphere(normal, rng);
                            const R lambert = std::max(dot(normal, rd), R(0)) / Pi();
                            throughput *= EvalDiffuse(normal, view, rd, diffuseColor, diffuseRoughness, diffuseModel) /
                                          std::max(lambert, R(1e-6));
                        }
                    }
                }
            }
            if (lightCount > 0U || hasMeshLights) {
                prevBsdfPdf = SurfacePdf(normal, tangent, view, rd, rx, ry, crx, cry, pSheen, pCoat, pSpec, pTl);
                prevDelta = false;
                envOnMiss = true;
                prevEnvPdf = lightCount > 0U
                    ? EnvLightPdf(hp, normal, rd, diffE, specE + LumGrey(coatContribution) + pSheen, tlE)
                    : R(0);
            }
            // Russian roulette (:1875-1879).
            if (bounce >= 2) {
                const R survival = clampR(throughput, R(0.1), R(0.95));
                if (Rand(rng) > survival) break;
                throughput /= survival;
            }
        }
        return radiance;
    }
};

struct OracleScene {
    explicit OracleScene(const hdcx_scene_view& v) : scene(v) {}
    Scene scene;
    bool fallbackInit = false;
    bool fallbackZUp = false;
};

thread_local std::string gError;

template <class R>
void RenderRows(const OracleScene& os, bool zUp, const hdcx_camera& cam, std::uint32_t width, std::uint32_t height,
                std::uint32_t sampleIndex, std::uint32_t maxBounces, std::uint32_t sampleCount,
                const std::vector<std::uint32_t>& rows, std::uint32_t colStep, int threads, double* out)
{
    const Integrator<R> integ{os.scene, os.scene.lightCount, zUp, os.scene.opaque};
    const int bounces = static_cast<int>(std::clamp(maxBounces, 1U, 12U));
    const std::uint32_t cols = (width + colStep - 1) / colStep;
    std::atomic<std::size_t> next{0};
    const std::size_t total = rows.size() * cols;
    auto worker = [&]() {
        for (;;) {
            const std::size_t item = next.fetch_add(1);
            if (item >= total) return;
            const std::uint32_t py = rows[item / cols];
            const std::uint32_t px = static_cast<std::uint32_t>(item % cols) * colStep;
            V3<R> batch(R(0));
            for (std::uint32_t so = 0; so < sampleCount; ++so) {
                const std::uint32_t sn = sampleIndex + so;
                V3<R> spectral(R(0));
                for (std::uint32_t lane = 0; lane < 3; ++lane) {
                    float wl;
                    const R L = integ.TraceLane(cam, width, height, px, py, sn, lane, bounces, wl);
                    spectral += Integrator<R>::Sensor(R(wl)) * L;
                }
                batch += spectral / R(3);
            }
            double* o = out + item * 3;
            o[0] = double(batch.x);
            o[1] = double(batch.y);
            o[2] = double(batch.z);
        }
    };
    std::vector<std::thread> pool;
    const int n = std::max(1, threads);
    for (int i = 0; i < n; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
}

} // namespace
} // namespace oracle

using oracle::OracleScene;

extern "C" {

const char* oracle_last_error(void) { return oracle::gError.c_str(); }

int oracle_scene_create(const hdcx_scene_view* view, void** out)
{
    try {
        if (!view || !out) throw std::invalid_argument("null argument");
        *out = new OracleScene(*view);
        return 0;
    } catch (const std::exception& e) {
        oracle::gError = e.what();
        return 1;
    }
}

void oracle_scene_destroy(void* s) { delete static_cast<OracleScene*>(s); }

// Sum over samples [sample_index, sample_index + sample_count) of the
// per-sample RGB (sum of 3 lanes' sensor x radiance / 3) for the pixels
// (row in rows[], column = k * col_step); out holds 3 doubles per pixel in
// row-major order of that subset. precision: 32 or 64.
int oracle_render(void* sp, const hdcx_camera* cam, std::uint32_t width, std::uint32_t height,
                  std::uint32_t sample_index, std::uint32_t max_bounces, std::uint32_t sample_count,
                  int precision, const std::uint32_t* rows, std::uint32_t row_count, std::uint32_t col_step,
                  int threads, double* out)
{
    try {
        auto* os = static_cast<OracleScene*>(sp);
        if (!os || !cam || !out || width == 0 || height == 0 || col_step == 0) throw std::invalid_argument("bad args");
        if (!os->fallbackInit) {
            // Fallback axis chosen from the first camera (:3237-3241).
            os->fallbackZUp = std::fabs(cam->vertical[2]) > std::fabs(cam->vertical[1]);
            os->fallbackInit = true;
        }
        std::vector<std::uint32_t> r(rows, rows + row_count);
        for (std::uint32_t y : r)
            if (y >= height) throw std::invalid_argument("row out of range");
        if (precision == 64)
            oracle::RenderRows<double>(*os, os->fallbackZUp, *cam, width, height, sample_index, max_bounces,
                                       std::max(sample_count, 1U), r, col_step, threads, out);
        else
            oracle::RenderRows<float>(*os, os->fallbackZUp, *cam, width, height, sample_index, max_bounces,
                                      std::max(sample_count, 1U), r, col_step, threads, out);
        return 0;
    } catch (const std::exception& e) {
        oracle::gError = e.what();
        return 1;
    }
}

// Closest hits for rays (8 floats each: o, tmin, d, tmax). out: t, prim,
// u, v as (float, uint32, float, float); t64 receives the FP64 t of the
// same primitive (NaN on a miss). brute_force != 0 tests every triangle.
int oracle_trace_closest(void* sp, const float* rays, std::uint32_t count, int brute_force, void* out_hits,
                         double* t64)
{
    try {
        auto* os = static_cast<OracleScene*>(sp);
        using oracle::V3;
        auto* hits = static_cast<std::uint32_t*>(out_hits);
        for (std::uint32_t i = 0; i < count; ++i) {
            const float* r = rays + static_cast<std::size_t>(i) * 8;
            const V3<float> o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
            const oracle::Hit h = brute_force ? oracle::ClosestHitBruteForce<float>(os->scene, o, d, r[3], r[7])
                                              : oracle::ClosestHit<float>(os->scene, o, d, r[3], r[7]);
            float t = h.found ? static_cast<float>(h.t) : 0.0F;
            float u = static_cast<float>(h.u), v = static_cast<float>(h.v);
            std::uint32_t prim = h.found ? h.primitive : 0xffffffffU;
            std::memcpy(&hits[i * 4 + 0], &t, 4);
            hits[i * 4 + 1] = prim;
            std::memcpy(&hits[i * 4 + 2], &u, 4);
            std::memcpy(&hits[i * 4 + 3], &v, 4);
            if (t64) {
                if (!h.found) {
                    t64[i] = NAN;
                } else {
                    const auto& s = os->scene;
                    const V3<double> od(r[0], r[1], r[2]), dd(r[4], r[5], r[6]);
                    const auto pre = oracle::Prepare<double>(od, dd);
                    double tt, uu, vv;
                    const bool ok = oracle::IntersectTriangle<double>(
                        pre, s.Pos<double>(s.indices[prim * 3]), s.Pos<double>(s.indices[prim * 3 + 1]),
                        s.Pos<double>(s.indices[prim * 3 + 2]), -INFINITY, INFINITY, tt, uu, vv);
                    t64[i] = ok ? tt : NAN;
                }
            }
        }
        return 0;
    } catch (const std::exception& e) {
        oracle::gError = e.what();
        return 1;
    }
}

// Occlusion in [tmin, tmax] against every triangle (opaque semantics).
int oracle_trace_occluded(void* sp, const float* rays, std::uint32_t count, std::uint8_t* out)
{
    auto* os = static_cast<OracleScene*>(sp);
    using oracle::V3;
    for (std::uint32_t i = 0; i < count; ++i) {
        const float* r = rays + static_cast<std::size_t>(i) * 8;
        const oracle::Hit h = oracle::ClosestHit<float>(os->scene, V3<float>(r[0], r[1], r[2]),
                                                        V3<float>(r[4], r[5], r[6]), r[3], r[7]);
        out[i] = h.found ? 1 : 0;
    }
    return 0;
}

// ---- Pure functions exposed for known-answer tests (FP64 unless noted) ----
std::uint32_t oracle_hash(std::uint32_t v) { return oracle::HashState(v); }
float oracle_random_float(std::uint32_t* state) { return oracle::RandomFloat(*state); }
void oracle_cie_xyz(double w, double* out)
{
    const auto v = oracle::Integrator<double>::CieXyz(w);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
void oracle_sensor(double w, double* out)
{
    const auto v = oracle::Integrator<double>::Sensor(w);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
double oracle_reconstruct_rgb(const double* rgb, double w)
{
    return oracle::Integrator<double>::ReconstructRgb(oracle::V3<double>(rgb[0], rgb[1], rgb[2]), w);
}
double oracle_illuminant(const double* rgb, double w)
{
    return oracle::Integrator<double>::Illuminant(oracle::V3<double>(rgb[0], rgb[1], rgb[2]), w);
}
double oracle_dispersed_ior(double ior, double scale, double abbe, double w)
{
    return oracle::Integrator<double>::DispersedIor(ior, scale, abbe, w);
}
// Diffuse BRDF x cos for the four models; n = +Z.
double oracle_eval_diffuse(const double* v, const double* l, double color, double rough, std::uint32_t model)
{
    using V = oracle::V3<double>;
    return oracle::Integrator<double>::EvalDiffuse(V(0, 0, 1), V(v[0], v[1], v[2]), V(l[0], l[1], l[2]), color, rough, model);
}
// GGX anisotropic reflection pdf, eval and sample weight; n = +Z, t = +X.
double oracle_ggx_pdf(const double* v, const double* l, double rx, double ry)
{
    using V = oracle::V3<double>;
    return oracle::Integrator<double>::GGXPdf(V(0, 0, 1), V(1, 0, 0), V(v[0], v[1], v[2]), V(l[0], l[1], l[2]), rx, ry);
}
double oracle_ggx_eval(const double* v, const double* l, double rx, double ry, double f0)
{
    using V = oracle::V3<double>;
    return oracle::Integrator<double>::EvalGGX(V(0, 0, 1), V(1, 0, 0), V(v[0], v[1], v[2]), V(l[0], l[1], l[2]), rx, ry, f0);
}
// Samples a GGX reflection for incident direction inc (pointing at the
// surface) with the FP32 random stream `state`; writes the outgoing
// direction and returns the sample weight.
double oracle_ggx_sample(const double* inc, double rx, double ry, double f0, std::uint32_t* state, double* out)
{
    using V = oracle::V3<double>;
    using I = oracle::Integrator<double>;
    const V n(0, 0, 1), t(1, 0, 0), i(inc[0], inc[1], inc[2]);
    const V o = I::SampleGGX(i, n, t, rx, ry, *state);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
    return I::GGXWeight(i, o, n, t, rx, ry, f0);
}
double oracle_sheen_albedo(double nv, double rough, int zeltner)
{
    using I = oracle::Integrator<double>;
    return zeltner ? I::SheenAlbedoZ(nv, rough) : I::SheenAlbedoCK(nv, rough);
}
double oracle_sheen_response(const double* v, const double* l, double rough, int zeltner)
{
    using V = oracle::V3<double>;
    return oracle::Integrator<double>::SheenResponse(V(0, 0, 1), V(v[0], v[1], v[2]), V(l[0], l[1], l[2]), rough, zeltner != 0);
}
double oracle_eon_albedo(double nd, double rough) { return oracle::Integrator<double>::EonAlbedo(nd, rough); }
void oracle_sample_hg(const double* fwd, double g, std::uint32_t* state, double* out)
{
    using V = oracle::V3<double>;
    const V o = oracle::Integrator<double>::SampleHG(V(fwd[0], fwd[1], fwd[2]), g, *state);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void oracle_cosine_hemisphere(const double* n, std::uint32_t* state, double* out)
{
    using V = oracle::V3<double>;
    const V o = oracle::Integrator<double>::CosineHemisphere(V(n[0], n[1], n[2]), *state);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void oracle_dome_uv(const double* d, std::uint32_t format, double* uv)
{
    using V = oracle::V3<double>;
    oracle::Integrator<double>::DomeUv(V(d[0], d[1], d[2]), format, uv[0], uv[1]);
}
void oracle_lat_long_direction(double u, double v, double* out)
{
    const auto d = oracle::Integrator<double>::LatLongDirection(u, v);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
// Samples the dome light `light` of the scene with the FP32 stream; returns
// the pdf and writes the direction; also returns the pdf the evaluation
// path (domeLightPdf) assigns to that direction in pdf_eval.
double oracle_sample_dome(void* sp, std::uint32_t light, std::uint32_t* state, double* dir, double* pdf_eval)
{
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    oracle::V3<double> d;
    double pdf;
    integ.SampleDomeLight(os->scene.lights[light], *state, d, pdf);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    if (pdf_eval) *pdf_eval = integ.DomePdf(os->scene.lights[light], d);
    return pdf;
}

// Transmission event at a boundary (Integrator::Transmit): rd is the
// incident direction, n the shading normal facing it, front whether the
// path enters the object; writes the new direction, returns 1 when the path
// crosses the boundary.
int oracle_transmission(const double* rd, const double* n, int front, double ior, int thin, std::uint32_t* state,
                        double* out)
{
    using V = oracle::V3<double>;
    V d(rd[0], rd[1], rd[2]);
    const bool through = oracle::Integrator<double>::Transmit(d, V(n[0], n[1], n[2]), front != 0, ior, thin != 0, *state);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
    return through ? 1 : 0;
}
// domeLightPdf (:583-598) of light `light` for world direction dir.
double oracle_dome_pdf(void* sp, std::uint32_t light, const double* dir)
{
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    return integ.DomePdf(os->scene.lights[light], oracle::V3<double>(dir[0], dir[1], dir[2]));
}
// sampleDome (:527-537): RGB radiance of dome light `light` toward dir.
void oracle_dome_radiance(void* sp, std::uint32_t light, const double* dir, double* out)
{
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    const auto c = integ.SampleDome(os->scene.lights[light], oracle::V3<double>(dir[0], dir[1], dir[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
// Bilinear REPEAT lookup of texture `texture` (decoded RGBA) at (u, v).
void oracle_texture(void* sp, std::uint32_t texture, double u, double v, double* out)
{
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    double a = 0.0;
    const auto c = integ.TexRgb(texture, u, v, &a);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = a;
}
// sampleMaterialTexture (:760-775) for an encoded texture slot (plain index,
// UDIM set 0x80000000 | mask << 8 | first, or 0xffffffff = missing).
void oracle_material_texture(void* sp, std::uint32_t slot, double u, double v, const double* fallback, double* out)
{
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    integ.SampleMat(slot, u, v, fallback, out);
}
// applyNormalMap (:807-829): p holds the triangle's three positions, uv its
// three corner UVs; gn is the (shading) normal the map perturbs.
void oracle_normal_map(void* sp, std::uint32_t texture, int flip_y, double scale, double u, double v, const double* gn,
                       const double* p, const double* uv, double* out)
{
    using V = oracle::V3<double>;
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    const V r = integ.NormalMap(texture, flip_y != 0, scale, u, v, V(gn[0], gn[1], gn[2]), V(p[0], p[1], p[2]),
                                V(p[3], p[4], p[5]), V(p[6], p[7], p[8]), uv[0], uv[1], uv[2], uv[3], uv[4], uv[5]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
// randomWalkSubsurface (:1153-1186) from entry point `entry` with outward
// normal out_n; returns 1 when the walk exits (exit origin, direction and
// attenuation written).
int oracle_random_walk(void* sp, const double* entry, const double* out_n, double mfp, double albedo, double aniso,
                       std::uint32_t* state, double* exit_o, double* exit_d, double* att)
{
    using V = oracle::V3<double>;
    auto* os = static_cast<OracleScene*>(sp);
    const oracle::Integrator<double> integ{os->scene, os->scene.lightCount, false, os->scene.opaque};
    V eo, ed;
    double a = 0.0;
    const bool exited = integ.RandomWalk(V(entry[0], entry[1], entry[2]), V(out_n[0], out_n[1], out_n[2]), mfp, albedo,
                                         aniso, *state, eo, ed, a);
    exit_o[0] = eo.x; exit_o[1] = eo.y; exit_o[2] = eo.z;
    exit_d[0] = ed.x; exit_d[1] = ed.y; exit_d[2] = ed.z;
    *att = a;
    return exited ? 1 : 0;
}

// The scene's emissive-triangle table (4 floats per triangle: cdf,
// pdf / area, area, 0), as given in the view or built by BuildEmissive.
void oracle_emissive_table(void* sp, float* out)
{
    const auto& sc = static_cast<OracleScene*>(sp)->scene;
    std::copy(sc.emissive.begin(), sc.emissive.end(), out);
}
// Importance-table size of light `light` (0 x 0: sampled uniformly) and,
// for each non-NULL pointer, its marginal CDF (height floats), conditional
// CDFs and pdf (width x height floats).
void oracle_dome_importance(void* sp, std::uint32_t light, std::uint32_t* width, std::uint32_t* height,
                            float* marginal, float* conditional, float* pdf)
{
    const auto& sc = static_cast<OracleScene*>(sp)->scene;
    const std::uint32_t ti = light < sc.lightCount && sc.lights[light].textureInfo.z == 0U
                                 ? sc.lights[light].textureInfo.w : oracle::kMissing;
    *width = *height = 0;
    if (ti == oracle::kMissing) return;
    const auto& t = sc.textures[ti];
    *width = static_cast<std::uint32_t>(t.width);
    *height = static_cast<std::uint32_t>(t.height);
    const std::size_t n = static_cast<std::size_t>(t.width) * t.height;
    for (std::size_t p = 0; p < n; ++p) {
        if (conditional) conditional[p] = t.texels[p * 4];
        if (pdf) pdf[p] = t.texels[p * 4 + 2];
    }
    if (marginal)
        for (int y = 0; y < t.height; ++y) marginal[y] = t.texels[static_cast<std::size_t>(y) * t.width * 4 + 1];
}

// UpscaleNearest (src/hydra/render_pass.cpp:45-68): destination pixel
// (x, y) copies source pixel (min(sw - 1, x * sw / w), min(sh - 1, y * sh / h)),
// unsigned 32-bit arithmetic; RGBA floats.
void oracle_upscale_nearest(const float* source, std::uint32_t sw, std::uint32_t sh, std::uint32_t w, std::uint32_t h,
                            float* out)
{
    for (std::uint32_t y = 0; y < h; ++y) {
        const std::uint32_t sy = std::min(sh - 1U, y * sh / h);
        for (std::uint32_t x = 0; x < w; ++x) {
            const std::uint32_t sx = std::min(sw - 1U, x * sw / w);
            std::memcpy(out + (static_cast<std::size_t>(y) * w + x) * 4, source + (static_cast<std::size_t>(sy) * sw + sx) * 4,
                        4 * sizeof(float));
        }
    }
}

} // extern "C"
