// ORACLE — test infrastructure only. Nothing in the product links this.
//
// Small vector helpers and the GLSL built-ins the reference shader uses,
// templated on the arithmetic type R (float = FP32 restatement, double =
// FP64 oracle). GLSL semantics restated here:
//   normalize(v) = v / length(v); reflect(I,N) = I - 2 dot(N,I) N;
//   refract(I,N,eta) per the GLSL 4.60 spec (zero vector on TIR);
//   mix(a,b,t) = a(1-t) + bt; smoothstep per spec; fract(x) = x - floor(x).
// GLSL pow(x, 2.0) is undefined for x < 0; the shader relies on it for
// signed bases (vulkan_path_tracer.cpp:121-123, :310-311, :458), so it is
// restated as x*x, the value every conforming implementation produces for
// x >= 0 and the evident intent for x < 0.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace oracle {

template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R s) : x(s), y(s), z(s) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(R s) const { return {x * s, y * s, z * s}; }
    V3 operator/(R s) const { return {x / s, y / s, z / s}; }
    V3 operator*(const V3& o) const { return {x * o.x, y * o.y, z * o.z}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
};

template <class R> inline V3<R> operator*(R s, const V3<R>& v) { return v * s; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(const V3<R>& v) { return std::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(const V3<R>& v) { return v / length(v); }
template <class R> inline R clampR(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <class R> inline R mixR(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> inline R sq(R x) { return x * x; }
template <class R> inline R fractR(R x) { return x - std::floor(x); }
template <class R> inline R smoothstepR(R e0, R e1, R x)
{
    const R t = clampR((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
template <class R> inline V3<R> reflect(const V3<R>& i, const V3<R>& n) { return i - n * (R(2) * dot(n, i)); }
template <class R> inline V3<R> refract(const V3<R>& i, const V3<R>& n, R eta)
{
    const R d = dot(n, i);
    const R k = R(1) - eta * eta * (R(1) - d * d);
    if (k < R(0)) return V3<R>(R(0));
    return i * eta - n * (eta * d + std::sqrt(k));
}
template <class R> inline V3<R> vmax(const V3<R>& a, R s)
{
    return {std::max(a.x, s), std::max(a.y, s), std::max(a.z, s)};
}

// ---- RNG: lowbias32 hash and its float conversion, always in FP32
// (vulkan_path_tracer.cpp:100-113). ----
inline std::uint32_t HashState(std::uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352dU;
    v ^= v >> 15;
    v *= 0x846ca68bU;
    return v ^ (v >> 16);
}

inline float RandomFloat(std::uint32_t& state)
{
    state = HashState(state);
    return static_cast<float>(state) * (1.0F / 4294967296.0F);
}

} // namespace oracle
