// Device shading library: spectral conversion, the compiled MaterialX
// closure (lobes, sampling, pdfs), light sampling and software texture
// filtering. Each function cites the reference shader lines
// (src/gpu/vulkan_path_tracer.cpp) whose semantics it implements. Transport
// is scalar per wavelength lane: every spectral vec3 of the reference shader
// has equal components, so one float carries it.
#pragma once

#include "dev_common.h"

namespace hdcx {

constexpr float kPi = 3.14159265359f;  // the reference's literal
constexpr float kTwoPi = 6.28318530718f;

// ---- Spectral conversion (:118-179) ----
__device__ __forceinline__ float reconstructRgb(V3 rgb, float w)
{
    // pow(x, 2.0) of the reference is x*x (its base may be negative).
    const float b0 = expf(-0.5f * sqf((w - 620.0f) / 30.0f));
    const float b1 = expf(-0.5f * sqf((w - 535.0f) / 25.0f));
    const float b2 = expf(-0.5f * sqf((w - 450.0f) / 20.0f));
    const V3 c = vmax(rgb, 0.0f);
    return (c.x * b0 + c.y * b1 + c.z * b2) / fmaxf(b0 + b1 + b2, 1e-5f);
}
__device__ __forceinline__ float reflectanceSpectrum(V3 rgb, float w) { return clampf(reconstructRgb(rgb, w), 0.0f, 1.0f); }
__device__ __forceinline__ float illuminantSpectrum(V3 rgb, float w)
{
    const float T = 6504.0f, c2 = 1.4387769e7f, ref = 560.0f;
    const float planck = powf(ref / w, 5.0f) * (expf(c2 / (ref * T)) - 1.0f) / fmaxf(expf(c2 / (w * T)) - 1.0f, 1e-5f);
    return fmaxf(reconstructRgb(rgb, w) * planck, 0.0f);
}

// CIE 1931 multi-lobe fit (Wyman, Sloan, Shirley, JCGT 2013) and the
// XYZ -> linear sRGB (IEC 61966-2-1) sensor, in FP64 (:146-179).
__device__ inline void spectralSensor64(double w, double out[3])
{
    const double tx1 = (w - 442.0) * (w < 442.0 ? 0.0624 : 0.0374);
    const double tx2 = (w - 599.8) * (w < 599.8 ? 0.0264 : 0.0323);
    const double tx3 = (w - 501.1) * (w < 501.1 ? 0.0490 : 0.0382);
    const double ty1 = (w - 568.8) * (w < 568.8 ? 0.0213 : 0.0247);
    const double ty2 = (w - 530.9) * (w < 530.9 ? 0.0613 : 0.0322);
    const double tz1 = (w - 437.0) * (w < 437.0 ? 0.0845 : 0.0278);
    const double tz2 = (w - 459.0) * (w < 459.0 ? 0.0385 : 0.0725);
    const double s = 400.0 / 106.856895;
    const double X = (0.362 * exp(-0.5 * tx1 * tx1) + 1.056 * exp(-0.5 * tx2 * tx2) - 0.065 * exp(-0.5 * tx3 * tx3)) * s;
    const double Y = (0.821 * exp(-0.5 * ty1 * ty1) + 0.286 * exp(-0.5 * ty2 * ty2)) * s;
    const double Z = (1.217 * exp(-0.5 * tz1 * tz1) + 0.681 * exp(-0.5 * tz2 * tz2)) * s;
    out[0] = 3.2404542 * X - 1.5371385 * Y - 0.4985314 * Z;
    out[1] = -0.9692660 * X + 1.8760108 * Y + 0.0415560 * Z;
    out[2] = 0.0556434 * X - 0.2040259 * Y + 1.0572252 * Z;
}

// luminance (:197-200) of a grey spectral vec3.
__device__ __forceinline__ float lumGrey(float s) { return s * 0.2126f + s * 0.7152f + s * 0.0722f; }
__device__ __forceinline__ float fresnelSchlick(float c, float f0)
{
    const float m = 1.0f - clampf(c, 0.0f, 1.0f);
    const float m2 = m * m;
    return f0 + (1.0f - f0) * (m2 * m2 * m);
}

__device__ __forceinline__ V3 anyTangent(V3 n)
{
    return normalize(fabsf(n.z) < 0.999f ? cross(n, v3(0.0f, 0.0f, 1.0f)) : cross(n, v3(0.0f, 1.0f, 0.0f)));
}
// cosineHemisphere (:183-195)
__device__ __forceinline__ V3 cosineHemisphere(V3 n, std::uint32_t& st)
{
    const float r1 = kTwoPi * randomFloat(st);
    const float r2 = randomFloat(st);
    const float r2s = sqrtf(r2);
    const V3 t = anyTangent(n);
    const V3 b = cross(n, t);
    float sn, cs;
    sincosf(r1, &sn, &cs);
    return normalize(t * (cs * r2s) + b * (sn * r2s) + n * sqrtf(1.0f - r2));
}
// uniformSphere (:1077-1083)
__device__ __forceinline__ V3 uniformSphere(std::uint32_t& st)
{
    const float y = 1.0f - 2.0f * randomFloat(st);
    const float radius = sqrtf(fmaxf(1.0f - y * y, 0.0f));
    const float phi = kTwoPi * randomFloat(st);
    float sn, cs;
    sincosf(phi, &sn, &cs);
    return v3(radius * cs, y, radius * sn);
}
// orthogonalTangent (:276-285)
__device__ __forceinline__ V3 orthogonalTangent(V3 n, V3 t)
{
    V3 r = t - n * dot(n, t);
    if (dot(r, r) < 1e-8f) r = fabsf(n.z) < 0.999f ? cross(n, v3(0.0f, 0.0f, 1.0f)) : cross(n, v3(0.0f, 1.0f, 0.0f));
    return normalize(r);
}

// ---- Anisotropic GGX (:287-383) ----
struct GgxFrame {
    V3 x, y;
    float ax, ay;
};
__device__ __forceinline__ GgxFrame ggxFrame(V3 n, V3 t, float rx, float ry)
{
    GgxFrame f;
    f.x = orthogonalTangent(n, t);
    f.y = cross(n, f.x);
    f.ax = clampf(rx * rx, 0.001f, 1.0f);
    f.ay = clampf(ry * ry, 0.001f, 1.0f);
    return f;
}
__device__ __forceinline__ float ggxD(const GgxFrame& f, V3 h, V3 n)
{
    const float lx = dot(h, f.x), ly = dot(h, f.y), lz = fmaxf(dot(h, n), 0.0f);
    const float den = lx * lx / (f.ax * f.ax) + ly * ly / (f.ay * f.ay) + lz * lz;
    return 1.0f / fmaxf(kPi * f.ax * f.ay * den * den, 1e-6f);
}
__device__ __forceinline__ float ggxG1(const GgxFrame& f, V3 dir, V3 n)
{
    const float nd = fmaxf(dot(n, dir), 0.0f);
    if (nd <= 0.0f) return 0.0f;
    const float proj = sqf(f.ax * dot(dir, f.x)) + sqf(f.ay * dot(dir, f.y));
    const float lambda = (-1.0f + sqrtf(1.0f + proj / fmaxf(nd * nd, 1e-8f))) * 0.5f;
    return 1.0f / (1.0f + lambda);
}
__device__ __forceinline__ float evalGgx(V3 n, V3 t, V3 v, V3 l, float rx, float ry, float f0)
{
    const V3 h = normalize(v + l);
    const float nv = fmaxf(dot(n, v), 0.0f), nl = fmaxf(dot(n, l), 0.0f), vh = fmaxf(dot(v, h), 0.0f);
    if (nv <= 0.0f || nl <= 0.0f || vh <= 0.0f) return 0.0f;
    const GgxFrame f = ggxFrame(n, t, rx, ry);
    const float D = ggxD(f, h, n);
    const float G = ggxG1(f, v, n) * ggxG1(f, l, n);
    return D * G * fresnelSchlick(vh, f0) / fmaxf(4.0f * nv * nl, 1e-5f);
}
__device__ __forceinline__ V3 sampleGgx(V3 inc, V3 n, V3 t, float rx, float ry, std::uint32_t& st)
{
    const GgxFrame f = ggxFrame(n, t, rx, ry);
    const float phi = kTwoPi * randomFloat(st);
    const float e = randomFloat(st);
    const float radius = sqrtf(e / fmaxf(1.0f - e, 1e-6f));
    float sn, cs;
    sincosf(phi, &sn, &cs);
    V3 h = normalize(f.x * (f.ax * radius * cs) + f.y * (f.ay * radius * sn) + n);
    if (dot(h, -inc) < 0.0f) h = -h;
    const V3 r = reflect3(inc, h);
    return dot(r, n) > 0.0f ? normalize(r) : normalize(reflect3(inc, n));
}
__device__ __forceinline__ float ggxSampleWeight(V3 inc, V3 out, V3 n, V3 t, float rx, float ry, float f0)
{
    const V3 v = normalize(-inc);
    const V3 h = normalize(v + out);
    const float nl = fmaxf(dot(n, out), 0.0f), nh = fmaxf(dot(n, h), 0.0f), vh = fmaxf(dot(v, h), 0.0f);
    if (nl <= 0.0f || nh <= 0.0f || vh <= 0.0f) return 0.0f;
    const GgxFrame f = ggxFrame(n, t, rx, ry);
    const float pdf = ggxD(f, h, n) * nh / fmaxf(4.0f * vh, 1e-6f);
    return evalGgx(n, t, v, out, rx, ry, f0) * nl / fmaxf(pdf, 1e-6f);
}
__device__ __forceinline__ float ggxPdf(V3 n, V3 t, V3 v, V3 l, float rx, float ry)
{
    if (dot(n, v) <= 0.0f || dot(n, l) <= 0.0f) return 0.0f;
    const V3 h = normalize(v + l);
    const float nh = fmaxf(dot(n, h), 0.0f), vh = fmaxf(dot(v, h), 0.0f);
    const GgxFrame f = ggxFrame(n, t, rx, ry);
    return ggxD(f, h, n) * nh / fmaxf(4.0f * vh, 1e-6f);
}
__device__ __forceinline__ float powerHeuristic(float a, float b)
{
    const float a2 = a * a, b2 = b * b;
    return a2 / fmaxf(a2 + b2, 1e-12f);
}

// ---- Sheen (:419-471) ----
__device__ __forceinline__ float sheenAlbedoCK(float nv, float rough)
{
    const float x = clampf(nv, 0.0f, 1.0f), y = clampf(rough, 0.01f, 1.0f);
    const float fx = 13.67300f + -68.78018f * x + 799.08825f * y + -905.00061f * x * y + 60.28956f * x * x +
                     1086.96473f * y * y;
    const float fy = 1.0f + 61.57746f * x + 442.78211f * y + 2597.49308f * x * y + 121.81241f * x * x +
                     3045.55075f * y * y;
    return clampf(fx / fmaxf(fy, 1e-6f), 0.0f, 1.0f);
}
__device__ __forceinline__ float sheenAlbedoZeltner(float nv, float rough)
{
    const float x = clampf(nv, 0.0f, 1.0f), y = clampf(rough, 0.01f, 1.0f);
    const float s = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
    const float m = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y * y))) / (0.046391f + y);
    const float o = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
    const float sd = fmaxf(s, 1e-6f);
    return clampf(expf(-0.5f * sqf((x - m) / sd)) / (sd * sqrtf(2.0f * kPi)) + o, 0.0f, 1.0f);
}
__device__ __forceinline__ float sheenResponse(V3 n, V3 v, V3 l, float rough, bool zeltner)
{
    const float nl = fmaxf(dot(n, l), 0.0f), nv = fmaxf(dot(n, v), 0.0f);
    if (!zeltner) {
        if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
        const V3 h = normalize(v + l);
        const float nh = fmaxf(dot(n, h), 0.0f);
        const float inv = 1.0f / fmaxf(rough, 0.005f);
        const float D = (2.0f + inv) * powf(fmaxf(1.0f - nh * nh, 0.0f), inv * 0.5f) / (2.0f * kPi);
        return D / fmaxf(4.0f * (nl + nv - nl * nv), 1e-6f) * nl;
    }
    return sheenAlbedoZeltner(nv, rough) * nl / kPi;
}

// ---- Diffuse models (:933-1009) ----
__device__ __forceinline__ float eonAlbedo(float nd, float rough)
{
    const float c1 = 0.2877934092f;
    const float c = clampf(nd, 0.0f, 1.0f);
    const float s = sqrtf(fmaxf(1.0f - c * c, 0.0f));
    const float theta = acosf(c);
    const float s3 = s * s * s;
    const float g = s * (theta - s * c) + (2.0f / 3.0f) * (s * (1.0f - s3) / fmaxf(c, 1e-5f) - s);
    const float a = 1.0f / (1.0f + c1 * rough);
    return clampf(a + rough * a * g / kPi, 0.0f, 1.0f);
}
__device__ __forceinline__ float evalDiffuse(V3 n, V3 v, V3 l, float color, float rough, std::uint32_t model)
{
    const float nl = fmaxf(dot(n, l), 0.0f), nv = fmaxf(dot(n, v), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
    if (model == 1U) {
        const float sg = clampf(rough, 0.0f, 1.0f), s2 = sg * sg;
        const float a = 1.0f - 0.5f * s2 / (s2 + 0.33f);
        const float b = 0.45f * s2 / (s2 + 0.09f);
        const V3 pl = l - n * nl, pv = v - n * nv;
        float az = 0.0f;
        if (dot(pl, pl) > 1e-8f && dot(pv, pv) > 1e-8f) az = fmaxf(dot(normalize(pl), normalize(pv)), 0.0f);
        const float sl = sqrtf(fmaxf(1.0f - nl * nl, 0.0f)), sv = sqrtf(fmaxf(1.0f - nv * nv, 0.0f));
        const float sa = fmaxf(sl, sv);
        const float tb = fminf(sl / fmaxf(nl, 1e-5f), sv / fmaxf(nv, 1e-5f));
        return nl * (a + b * az * sa * tb) / kPi;
    }
    if (model == 2U) {
        const V3 h = normalize(v + l);
        const float lh = fmaxf(dot(l, h), 0.0f);
        const float grazing = 0.5f + 2.0f * clampf(rough, 0.0f, 1.0f) * lh * lh;
        const float ml = 1.0f - nl, mv = 1.0f - nv;
        const float ls = 1.0f + (grazing - 1.0f) * (ml * ml * ml * ml * ml);
        const float vs = 1.0f + (grazing - 1.0f) * (mv * mv * mv * mv * mv);
        return nl * ls * vs / kPi;
    }
    if (model == 3U) {
        const float c1 = 0.2877934092f, c2 = 0.0726760455f;
        const float sg = clampf(rough, 0.0f, 1.0f);
        const float a = 1.0f / (1.0f + c1 * sg);
        const float sv = dot(l, v) - nl * nv;
        const float st = sv > 0.0f ? sv / fmaxf(nl, nv) : sv;
        const float single = a * (1.0f + sg * st);
        const float avg = (1.0f + c2 * sg) / (1.0f + c1 * sg);
        const float dirf = (1.0f - eonAlbedo(nv, sg)) * (1.0f - eonAlbedo(nl, sg)) / fmaxf(1.0f - avg, 1e-5f);
        const float rho = clampf(color, 0.0f, 0.9999f);
        const float ms = rho * avg / fmaxf(1.0f - rho * (1.0f - avg), 1e-5f);
        return nl / kPi * (single + ms * dirf);
    }
    return nl / kPi;
}

// The resolved closure at a hit (material fetch :1296-1464).
struct Closure {
    V3 normal, tangent, view;
    float diffuseColor, transmission, subsurface, diffuseWeight, diffuseRoughness;
    std::uint32_t diffuseModel;
    float translucentWeight, translucentColor, metalness, rx, ry, f0;
    float coat, crx, cry, coatF0;
    float sheen, sheenColor, sheenRoughness;
    bool sheenZeltner;
    float sheenAlbedo, pSheen, coatContribution, pCoat, pSpec, pTranslucent;
    float specE, diffE, translucentE, diffuseContribution, translucentContribution;
};

// evaluateDirectSurface (:1013-1069): BSDF x cos toward l. kFull = false:
// the closure has no sheen (c.sheen == 0), whose layer is then left out
// (it is 0 x a finite response).
template <bool kFull = true>
__device__ inline float evalDirect(const Closure& c, V3 l, float diffuseScale, float specularScale)
{
    const V3 n = c.normal, v = c.view;
    const float snl = dot(n, l);
    const float nl = fmaxf(snl, 0.0f), bnl = fmaxf(-snl, 0.0f);
    const float wrapped = fmaxf((snl + 0.5f * c.subsurface) / (1.0f + 0.5f * c.subsurface), 0.0f);
    if (wrapped <= 0.0f && nl <= 0.0f && bnl * c.translucentWeight <= 0.0f) return 0.0f;
    const V3 h = nl > 0.0f ? normalize(v + l) : n;
    const float vh = fmaxf(dot(v, h), 0.0f);
    const float baseF = fresnelSchlick(vh, c.f0);
    const float dresp = mixf(evalDiffuse(n, v, l, c.diffuseColor, c.diffuseRoughness, c.diffuseModel), wrapped / kPi,
                             c.subsurface);
    const float diffuse = diffuseScale * (1.0f - baseF) * c.diffuseWeight * (1.0f - c.translucentWeight) *
                          (1.0f - c.metalness) * c.diffuseColor * dresp * (1.0f - c.transmission);
    const float nv = fmaxf(dot(n, v), 0.0f);
    const float viewF = fresnelSchlick(nv, c.f0);
    const float translucent = diffuseScale * (1.0f - viewF) * c.translucentWeight * (1.0f - c.metalness) *
                              c.translucentColor * (bnl / kPi) * (1.0f - c.transmission);
    const float specular = specularScale * nl * evalGgx(n, c.tangent, v, l, c.rx, c.ry, c.f0);
    float coatF = 0.0f, coatSpec = 0.0f;
    if (c.coat > 0.0f) {
        coatF = fresnelSchlick(vh, c.coatF0);
        coatSpec = specularScale * c.coat * nl * evalGgx(n, c.tangent, v, l, c.crx, c.cry, c.coatF0);
    }
    const float baseLayer = (1.0f - c.coat * coatF) * (diffuse + translucent + specular) + coatSpec;
    const float sheenLayer =
        kFull ? specularScale * c.sheen * c.sheenColor * sheenResponse(n, v, l, c.sheenRoughness, c.sheenZeltner) : 0.0f;
    return baseLayer * (1.0f - clampf(c.sheen * c.sheenAlbedo, 0.0f, 1.0f)) + sheenLayer;
}

// surfaceReflectionPdf (:392-415)
// kFull = false: the closure has no coat (c.pCoat == 0), whose pdf term is
// then left out (it is 0 x a finite pdf).
template <bool kFull = true>
__device__ inline float surfacePdf(const Closure& c, V3 l)
{
    const float refl = fmaxf(dot(c.normal, l), 0.0f) / kPi;
    const float trans = fmaxf(-dot(c.normal, l), 0.0f) / kPi;
    const float coatPdf = kFull ? ggxPdf(c.normal, c.tangent, c.view, l, c.crx, c.cry) : 0.0f;
    const float specPdf = ggxPdf(c.normal, c.tangent, c.view, l, c.rx, c.ry);
    const float diffPdf = mixf(refl, trans, c.pTranslucent);
    const float base = c.pCoat * coatPdf + (1.0f - c.pCoat) * (c.pSpec * specPdf + (1.0f - c.pSpec) * diffPdf);
    return c.pSheen * refl + (1.0f - c.pSheen) * base;
}

// ---- Software texture filtering: bilinear, REPEAT, no mips (sampler
// :2401-2423); RGBA8 sRGB texels decoded before filtering. ----
__device__ __forceinline__ float srgbDecode(float c)
{
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}
__device__ __forceinline__ float4 fetchTexel(const DeviceScene& s, const TexDesc& d, int x, int y)
{
    const std::size_t i = static_cast<std::size_t>(y) * d.width + x;
    if (d.format == HDCX_TEXTURE_RGBA32F) {
        return reinterpret_cast<const float4*>(s.texels + d.offset)[i];
    }
    const uchar4 b = reinterpret_cast<const uchar4*>(s.texels + d.offset)[i];
    float4 f = make_float4(b.x / 255.0f, b.y / 255.0f, b.z / 255.0f, b.w / 255.0f);
    if (d.format == HDCX_TEXTURE_RGBA8_SRGB) {
        f.x = srgbDecode(f.x);
        f.y = srgbDecode(f.y);
        f.z = srgbDecode(f.z);
    }
    return f;
}
__device__ __forceinline__ int wrapIndex(int i, int n)
{
    const int m = i % n;
    return m < 0 ? m + n : m;
}
__device__ inline float4 sampleTexture(const DeviceScene& s, std::uint32_t index, float u, float v)
{
    const TexDesc d = s.textures[index];
    float x = u * static_cast<float>(d.width) - 0.5f, y = v * static_cast<float>(d.height) - 0.5f;
    if (!isfinite(x) || !isfinite(y)) { x = 0.0f; y = 0.0f; }
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    // Reduce before converting so huge coordinates stay in range.
    const float rx = fx - floorf(fx / d.width) * d.width, ry = fy - floorf(fy / d.height) * d.height;
    const int ix = static_cast<int>(rx), iy = static_cast<int>(ry);
    const int x0 = wrapIndex(ix, d.width), x1 = wrapIndex(ix + 1, d.width);
    const int y0 = wrapIndex(iy, d.height), y1 = wrapIndex(iy + 1, d.height);
    const float4 t00 = fetchTexel(s, d, x0, y0), t10 = fetchTexel(s, d, x1, y0);
    const float4 t01 = fetchTexel(s, d, x0, y1), t11 = fetchTexel(s, d, x1, y1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return make_float4((t00.x * bx + t10.x * ax) * by + (t01.x * bx + t11.x * ax) * ay,
                       (t00.y * bx + t10.y * ax) * by + (t01.y * bx + t11.y * ax) * ay,
                       (t00.z * bx + t10.z * ax) * by + (t01.z * bx + t11.z * ax) * ay,
                       (t00.w * bx + t10.w * ax) * by + (t01.w * bx + t11.w * ax) * ay);
}
// sampleMaterialTexture (:760-775) including UDIM tile selection.
__device__ inline float4 sampleMaterialTexture(const DeviceScene& s, std::uint32_t ti, float u, float v, float4 fb)
{
    if (ti == kMissingTexture) return fb;
    if ((ti & kUdimTextureFlag) != 0U) {
        const float tu = floorf(u), tv = floorf(v);
        if (!(fabsf(tu) < 1e6f && fabsf(tv) < 1e6f)) return fb;
        const int off = static_cast<int>(tu) + static_cast<int>(tv) * 10;
        if (off < 0 || off >= 23) return fb;
        const std::uint32_t present = (ti >> 8U) & kUdimTileMask;
        const std::uint32_t bit = 1U << static_cast<std::uint32_t>(off);
        if ((present & bit) == 0U) return fb;
        const std::uint32_t di = (ti & 0xffU) + static_cast<std::uint32_t>(__popc(present & (bit - 1U)));
        return sampleTexture(s, di, fractf(u), fractf(v));
    }
    return sampleTexture(s, ti, u, v);
}
__device__ __forceinline__ float matScalar(const DeviceScene& s, std::uint32_t ti, float u, float v, float fb)
{
    if (ti == kMissingTexture) return fb;
    return sampleMaterialTexture(s, ti, u, v, make_float4(fb, fb, fb, fb)).x;
}
__device__ __forceinline__ V3 matRgb(const DeviceScene& s, std::uint32_t ti, float u, float v, V3 fb)
{
    if (ti == kMissingTexture) return fb;
    const float4 r = sampleMaterialTexture(s, ti, u, v, make_float4(fb.x, fb.y, fb.z, 1.0f));
    return v3(r.x, r.y, r.z);
}

// ---- Dome lights (:473-650) ----
__device__ inline void domeUv(V3 d, std::uint32_t format, float& u, float& v)
{
    if (format == 2U) {
        const float den = 2.0f * sqrtf(fmaxf(d.x * d.x + d.y * d.y + (d.z + 1.0f) * (d.z + 1.0f), 1e-8f));
        u = d.x / den + 0.5f;
        v = d.y / den + 0.5f;
        return;
    }
    if (format == 3U) {
        const float radius = acosf(clampf(-d.z, -1.0f, 1.0f)) / kPi;
        const float len = sqrtf(d.x * d.x + d.y * d.y);
        float rx = 0.0f, ry = 0.0f;
        if (len > 1e-6f) { rx = d.x / len; ry = d.y / len; }
        u = 0.5f + rx * (0.5f * radius);
        v = 0.5f + ry * (0.5f * radius);
        return;
    }
    if (format == 4U) {
        const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        float fu, fv, tu, tv;
        if (ax >= ay && ax >= az) {
            if (d.x >= 0.0f) { fu = -d.z / ax; fv = d.y / ax; tu = 2.0f; tv = 1.0f; }
            else { fu = d.z / ax; fv = d.y / ax; tu = 0.0f; tv = 1.0f; }
        } else if (ay >= az) {
            if (d.y >= 0.0f) { fu = d.x / ay; fv = -d.z / ay; tu = 1.0f; tv = 0.0f; }
            else { fu = d.x / ay; fv = d.z / ay; tu = 1.0f; tv = 2.0f; }
        } else if (d.z >= 0.0f) { fu = d.x / az; fv = d.y / az; tu = 1.0f; tv = 1.0f; }
        else { fu = -d.x / az; fv = d.y / az; tu = 1.0f; tv = 3.0f; }
        u = (tu + fu * 0.5f + 0.5f) / 3.0f;
        v = (tv + fv * 0.5f + 0.5f) / 4.0f;
        return;
    }
    u = fractf(atan2f(d.x, -d.z) / (2.0f * kPi) + 1.0f);
    v = 1.0f - acosf(clampf(d.y, -1.0f, 1.0f)) / kPi;
}
__device__ __forceinline__ V3 domeLocal(const GpuLightRecord& L, V3 wd)
{
    return normalize(v3(dot(wd, v3(L.basisX)), dot(wd, v3(L.basisY)), dot(wd, v3(L.basisZ))));
}
__device__ inline V3 sampleDome(const DeviceScene& s, const GpuLightRecord& L, V3 wd)
{
    V3 tex = v3(1.0f);
    if (L.textureInfo.x != kMissingTexture) {
        float u, v;
        domeUv(domeLocal(L, wd), L.textureInfo.y, u, v);
        const float4 t = sampleTexture(s, L.textureInfo.x, u, v);
        tex = v3(t.x, t.y, t.z);
    }
    return v3(L.colorIntensity) * L.colorIntensity.w * tex;
}
__device__ inline V3 environmentRgb(const DeviceScene& s, V3 d)
{
    if (s.lightCount == 0U) {
        if (s.hasEmissive) return v3(0.0f);
        const V3 up = s.zUp ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
        const float t = 0.5f * (dot(d, up) + 1.0f);
        return v3(0.035f, 0.045f, 0.065f) * (1.0f - t) + v3(0.38f, 0.55f, 0.85f) * t;
    }
    V3 r = v3(0.0f);
    for (std::uint32_t i = 0; i < s.lightCount; ++i) {
        const GpuLightRecord& L = s.lights[i];
        if (L.textureInfo.z == 0U) r = r + sampleDome(s, L, d);
    }
    return r;
}
__device__ __forceinline__ V3 latLongDirection(float u, float v)
{
    const float theta = kPi * (1.0f - v), phi = 2.0f * kPi * u;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    return v3(st * sp, ct, -st * cp);
}
// domeLightPdf (:583-598) from the compact importance tables.
__device__ inline float domeLightPdf(const DeviceScene& s, std::uint32_t li, const GpuLightRecord& L, V3 wd)
{
    if (L.textureInfo.w == kMissingTexture) return 1.0f / 12.5663706144f;
    float u, v;
    domeUv(domeLocal(L, wd), L.textureInfo.y, u, v);
    const ImportanceDesc d = s.importance[li];
    const int tx = min(max(static_cast<int>(u * static_cast<float>(d.width)), 0), static_cast<int>(d.width) - 1);
    const int ty = min(max(static_cast<int>(v * static_cast<float>(d.height)), 0), static_cast<int>(d.height) - 1);
    return fmaxf(s.importancePool[d.pdf + static_cast<std::size_t>(ty) * d.width + tx], 0.0f);
}
// First index i of cdf[0, n) with u <= cdf[i] (n - 1 when there is none),
// the binary search of sampleDomeLight. With a guide table the search runs
// only between guide[g] and guide[g + 1], g = floor(u * 2^shift) (exact:
// a power of two), which hold the answer when the CDF is monotone: every
// index below guide[g] has a value < g / 2^shift <= u, and cdf[guide[g + 1]]
// >= (g + 1) / 2^shift > u. Same index, a few probes instead of log2(n).
__device__ __forceinline__ int cdfSearch(const float* cdf, int n, float u, const std::uint32_t* guide, std::uint32_t shift)
{
    int lo = 0, hi = n - 1;
    if (guide) {
        const std::uint32_t buckets = 1U << shift;
        const std::uint32_t g = min(static_cast<std::uint32_t>(u * static_cast<float>(buckets)), buckets - 1U);
        lo = static_cast<int>(guide[g]);
        hi = static_cast<int>(guide[g + 1U]);
    }
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (u <= cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// sampleDomeLight (:600-650): row by marginal CDF, column by conditional
// CDF, cos(theta) stratified within the row.
__device__ inline void sampleDomeLight(const DeviceScene& s, std::uint32_t li, const GpuLightRecord& L,
                                       std::uint32_t& rng, V3& wd, float& pdf)
{
    if (L.textureInfo.w == kMissingTexture) {
        wd = uniformSphere(rng);
        pdf = 1.0f / 12.5663706144f;
        return;
    }
    const ImportanceDesc d = s.importance[li];
    const std::uint32_t* guides = reinterpret_cast<const std::uint32_t*>(s.importancePool);
    const float rowTarget = randomFloat(rng);
    const int row = cdfSearch(s.importancePool + d.marginal, static_cast<int>(d.height), rowTarget,
                              d.rowGuide ? guides + d.rowGuide : nullptr, d.rowShift);
    const float colTarget = randomFloat(rng);
    const int col = cdfSearch(s.importancePool + d.conditional + static_cast<std::size_t>(row) * d.width,
                              static_cast<int>(d.width), colTarget,
                              d.colGuide ? guides + d.colGuide + static_cast<std::size_t>(row) * ((1U << d.colShift) + 1U) : nullptr,
                              d.colShift);
    const float v0 = static_cast<float>(row) / static_cast<float>(d.height);
    const float v1 = static_cast<float>(row + 1) / static_cast<float>(d.height);
    const float c0 = cosf(kPi * (1.0f - v0)), c1 = cosf(kPi * (1.0f - v1));
    const float ct = mixf(c0, c1, randomFloat(rng));
    const float v = 1.0f - acosf(clampf(ct, -1.0f, 1.0f)) / kPi;
    const float u = (static_cast<float>(col) + randomFloat(rng)) / static_cast<float>(d.width);
    const V3 local = latLongDirection(u, v);
    wd = normalize(v3(L.basisX) * local.x + v3(L.basisY) * local.y + v3(L.basisZ) * local.z);
    pdf = fmaxf(s.importancePool[d.pdf + static_cast<std::size_t>(row) * d.width + col], 0.0f);
}

// ---- Light selection (:652-758) ----
__device__ inline float lightSelectionWeight(const GpuLightRecord& L, V3 pos, V3 n, float de, float se, float te)
{
    const float power = fmaxf(L.geometry.w, 0.0f);
    if (power <= 0.0f) return 0.0f;
    const float dr = fmaxf(de, 0.0f) * L.controls.x;
    const float sr = fmaxf(se, 0.0f) * L.controls.y;
    const float tr = fmaxf(te, 0.0f) * L.controls.x;
    const float minimum = power * 0.01f;
    const std::uint32_t type = L.textureInfo.z;
    if (type == 0U) return fmaxf(minimum, power * (0.5f * (dr + tr) + sr));
    V3 dir;
    float geom = 1.0f;
    if (type == 5U) {
        dir = normalize(v3(L.basisZ));
    } else {
        const V3 toL = v3(L.position) - pos;
        const float d2 = fmaxf(dot(toL, toL), 1e-4f);
        dir = toL * (1.0f / sqrtf(d2));
        geom = 1.0f / d2;
        if (type == 1U || type == 2U) {
            const V3 ln = -normalize(v3(L.basisZ));
            geom *= fmaxf(dot(ln, -dir), 0.0f);
        } else if (type == 4U) {
            geom *= 0.5f;
        }
    }
    const float nl = dot(n, dir);
    const float response = fmaxf(nl, 0.0f) * (dr + sr) + fmaxf(-nl, 0.0f) * tr;
    return fmaxf(minimum, power * geom * response);
}
__device__ inline float totalLightWeight(const DeviceScene& s, V3 pos, V3 n, float de, float se, float te)
{
    float total = 0.0f;
    for (std::uint32_t i = 0; i < s.lightCount; ++i) total += lightSelectionWeight(s.lights[i], pos, n, de, se, te);
    return total;
}
// total: totalLightWeight at the same point and lobe energies (one sum per
// shading point, shared by light selection and both MIS pdfs).
__device__ inline void sampleAnalyticLight(const DeviceScene& s, V3 pos, V3 n, float de, float se, float te, float total,
                                           std::uint32_t& rng, std::uint32_t& index, float& pdf)
{
    if (total <= 0.0f) {
        index = min(static_cast<std::uint32_t>(randomFloat(rng) * static_cast<float>(s.lightCount)), s.lightCount - 1U);
        pdf = 1.0f / static_cast<float>(s.lightCount);
        return;
    }
    const float target = randomFloat(rng) * total;
    float acc = 0.0f;
    index = s.lightCount - 1U;
    for (std::uint32_t i = 0; i < s.lightCount; ++i) {
        acc += lightSelectionWeight(s.lights[i], pos, n, de, se, te);
        if (target <= acc) { index = i; break; }
    }
    pdf = lightSelectionWeight(s.lights[index], pos, n, de, se, te) / total;
}
__device__ inline float environmentLightPdf(const DeviceScene& s, V3 pos, V3 n, V3 dir, float de, float se, float te,
                                           float total)
{
    if (s.lightCount == 0U) return 0.0f;
    if (total <= 0.0f) return 0.0f;
    float pdf = 0.0f;
    for (std::uint32_t i = 0; i < s.lightCount; ++i) {
        const GpuLightRecord& L = s.lights[i];
        if (L.textureInfo.z == 0U) pdf += lightSelectionWeight(L, pos, n, de, se, te) / total * domeLightPdf(s, i, L, dir);
    }
    return pdf;
}
// shapingConeFactor (:1085-1094)
__device__ __forceinline__ float shapingCone(const GpuLightRecord& L, float cosine)
{
    const float angle = acosf(clampf(cosine, -1.0f, 1.0f));
    const float outer = clampf(L.shaping.y, 0.0f, 180.0f) * (kPi / 180.0f);
    const float inner = outer * (1.0f - clampf(L.shaping.z, 0.0f, 1.0f));
    if (outer >= 3.1415925f) return 1.0f;
    if (outer <= inner + 1e-6f) return angle <= outer ? 1.0f : 0.0f;
    const float t = clampf((angle - inner) / (outer - inner), 0.0f, 1.0f);
    return 1.0f - t * t * (3.0f - 2.0f * t);
}
// dispersedIor (:1096-1110): Cauchy fit through n_d with Abbe number.
__device__ __forceinline__ float dispersedIor(float iorD, float scale, float abbe, float w)
{
    if (scale <= 0.0f || abbe <= 1.0f) return iorD;
    const float lf = 486.13f, ld = 587.56f, lc = 656.27f;
    const float rd = 1.0f / (lf * lf) - 1.0f / (lc * lc);
    const float b = (iorD - 1.0f) / fmaxf(abbe * rd, 1e-8f);
    const float a = iorD - b / (ld * ld);
    return mixf(iorD, a + b / (w * w), clampf(scale, 0.0f, 1.0f));
}
// sampleHenyeyGreenstein (:1112-1128)
__device__ inline V3 sampleHG(V3 fwd, float aniso, std::uint32_t& st)
{
    const float g = clampf(aniso, -0.95f, 0.95f);
    const float u = randomFloat(st);
    const float cosine = fabsf(g) < 1e-3f ? 1.0f - 2.0f * u
                                           : (1.0f + g * g - sqf((1.0f - g * g) / (1.0f - g + 2.0f * g * u))) / (2.0f * g);
    const float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.0f));
    const float phi = kTwoPi * randomFloat(st);
    float sn, cs;
    sincosf(phi, &sn, &cs);
    const V3 w = normalize(fwd);
    const V3 t = anyTangent(w);
    const V3 b = cross(w, t);
    return normalize(t * (cs * sine) + b * (sn * sine) + w * cosine);
}
// offsetSurfaceOrigin (:1071-1075)
__device__ __forceinline__ V3 offsetOrigin(V3 p, V3 gn, V3 d)
{
    const float side = dot(gn, d) >= 0.0f ? 1.0f : -1.0f;
    return p + gn * (0.002f * side);
}

} // namespace hdcx
