// Instance transforms, written once for the device (on-the-fly instanced
// triangles) and the host (the flattened scene the oracle renders), so that
// both produce the same FP32 world-space vertices and normals bit for bit:
// fixed operation order, no contraction, IEEE division and square root.
#pragma once

#include "lbvh_core.h"

#include <cmath>
#include <cstdint>

namespace hdcx {

// Device copy of one instance (144 B): object-to-world rows, world-to-object
// rows and the normal matrix (inverse transpose) rows, the last two
// computed in FP64 on the host and rounded once.
struct alignas(16) InstanceRecord {
    float toWorld[12];
    float toObject[12];
    float normal[9];
    std::uint32_t prototype;
    std::uint32_t material;   // 0xffffffff: the prototype triangles' own
    std::uint32_t base;       // flattened id of the instance's first triangle
};
static_assert(sizeof(InstanceRecord) == 144, "instance record");

// Per prototype: its triangle range in the scene arrays and its BLAS.
struct alignas(16) PrototypeRecord {
    std::uint32_t firstTriangle, triangleCount;
    std::uint32_t root;       // BLAS root node
    std::uint32_t pad;
};

// Row r of a 3x4 transform applied to a point: ((m0 x + m1 y) + m2 z) + m3.
HDCX_HD inline void TransformPoint(const float* m, float x, float y, float z, float out[3])
{
    HDCX_NO_CONTRACT
    for (int r = 0; r < 3; ++r) out[r] = ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3];
}

// Normal matrix applied to a corner normal, then normalised; a zero normal
// (meaning "use the geometric normal") stays zero. The square root and the
// divisions are the correctly rounded FP32 ones in every build (host
// flattening, and device code built with hardware-rate FP32 division):
// each is evaluated in FP64 and rounded once to FP32, which gives the
// correctly rounded FP32 result for sqrt and division of FP32 operands.
HDCX_HD inline void TransformNormal(const float* nm, float x, float y, float z, float out[3])
{
    HDCX_NO_CONTRACT
    float v[3];
    for (int r = 0; r < 3; ++r) v[r] = (nm[3 * r] * x + nm[3 * r + 1] * y) + nm[3 * r + 2] * z;
    const float l2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (!(l2 > 0.0f)) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    const float l = static_cast<float>(sqrt(static_cast<double>(l2)));
    for (int r = 0; r < 3; ++r) out[r] = static_cast<float>(static_cast<double>(v[r]) / static_cast<double>(l));
}

// Host: inverse of the 3x4 affine transform and the normal matrix, in FP64,
// rounded once. Returns false for a singular linear part.
inline bool InvertInstance(const float* m, float toObject[12], float normal[9])
{
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double inv[9] = {A / det, -(b * i - c * h) / det, (b * f - c * e) / det,
                           B / det, (a * i - c * g) / det, -(a * f - c * d) / det,
                           C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
    const double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) toObject[4 * r + k] = static_cast<float>(inv[3 * r + k]);
        toObject[4 * r + 3] = static_cast<float>(-(inv[3 * r] * t[0] + inv[3 * r + 1] * t[1] + inv[3 * r + 2] * t[2]));
    }
    // Normal matrix = transpose of the inverse.
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) normal[3 * r + k] = static_cast<float>(inv[3 * k + r]);
    return true;
}

} // namespace hdcx
