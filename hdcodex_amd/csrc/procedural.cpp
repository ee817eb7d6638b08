// Seeded procedural scenes. The reference builds its scenes from Hydra; these
// generators stand in for that producer. They emit SceneSnapshot objects in
// publish order (ids sorted) and run them through the regular packer.
#include "procedural.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace hdcx {
namespace {

constexpr double kPi = 3.14159265358979323846;

void AddQuad(SceneMesh& m, const double p[4][3])
{
    const std::uint32_t base = static_cast<std::uint32_t>(m.positions.size() / 3);
    for (int i = 0; i < 4; ++i)
        for (int c = 0; c < 3; ++c) m.positions.push_back(static_cast<float>(p[i][c]));
    m.indices.insert(m.indices.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
}

// Axis-aligned box of half-extents h rotated by `angle` about +Y around its
// centre c; 6 faces, 12 triangles.
SceneMesh RotatedBox(const std::string& id, const double c[3], const double h[3], double angle)
{
    SceneMesh m;
    m.id = id;
    const double ca = std::cos(angle), sa = std::sin(angle);
    auto corner = [&](int sx, int sy, int sz, double out[3]) {
        const double lx = sx * h[0], ly = sy * h[1], lz = sz * h[2];
        out[0] = c[0] + ca * lx + sa * lz;
        out[1] = c[1] + ly;
        out[2] = c[2] - sa * lx + ca * lz;
    };
    const int faces[6][4][3] = {
        {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}},      // +z
        {{1, -1, -1}, {-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}},  // -z
        {{1, -1, 1}, {1, -1, -1}, {1, 1, -1}, {1, 1, 1}},      // +x
        {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, -1}},  // -x
        {{-1, 1, 1}, {1, 1, 1}, {1, 1, -1}, {-1, 1, -1}},      // +y
        {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}},  // -y
    };
    for (const auto& face : faces) {
        double q[4][3];
        for (int i = 0; i < 4; ++i) corner(face[i][0], face[i][1], face[i][2], q[i]);
        AddQuad(m, q);
    }
    return m;
}

SceneMaterial Lambert(const std::string& id, SplitMix64& rng)
{
    SceneMaterial m;
    m.id = id;
    m.baseColor = {rng.Range(0.1, 0.9), rng.Range(0.1, 0.9), rng.Range(0.1, 0.9)};
    m.diffuseModel = SceneMaterial::DiffuseModel::Lambert;
    m.specularWeight = 0.0F;
    return m;
}

template <class T>
void SortById(std::vector<T>& v)
{
    std::sort(v.begin(), v.end(), [](const T& a, const T& b) { return a.id < b.id; });
}

void SortSnapshot(SceneSnapshot& s)
{
    SortById(s.meshes);
    SortById(s.materials);
    SortById(s.textures);
    SortById(s.lights);
}

// ---- Cornell box: 5 wall quads + 2 rotated boxes = 34 triangles, one
// 0.25 x 0.25 rect light under the ceiling facing -Y, Lambert albedos
// U[0.1, 0.9] per RGB channel per object. ----
ProceduralScene MakeCornell(std::uint64_t seed)
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    const char* wallNames[5] = {"back", "ceiling", "floor", "left", "right"};
    const double walls[5][4][3] = {
        {{-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}},   // back (z = -1)
        {{-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}},     // ceiling (y = 2)
        {{-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, -1}},     // floor (y = 0)
        {{-1, 0, 1}, {-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}},   // left (x = -1)
        {{1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}},       // right (x = +1)
    };
    for (int i = 0; i < 5; ++i) {
        SceneMesh m;
        m.id = std::string("/cornell/wall_") + wallNames[i];
        AddQuad(m, walls[i]);
        m.materialId = std::string("/cornell/materials/wall_") + wallNames[i];
        s.meshes.push_back(m);
        SplitMix64 r{StreamSeed(seed, 100 + i)};
        s.materials.push_back(Lambert(m.materialId, r));
    }
    const double shortC[3] = {0.35, 0.3, 0.3}, shortH[3] = {0.3, 0.3, 0.3};
    const double tallC[3] = {-0.35, 0.6, -0.3}, tallH[3] = {0.3, 0.6, 0.3};
    SceneMesh shortBox = RotatedBox("/cornell/box_short", shortC, shortH, -18.0 * kPi / 180.0);
    SceneMesh tallBox = RotatedBox("/cornell/box_tall", tallC, tallH, 20.0 * kPi / 180.0);
    shortBox.materialId = "/cornell/materials/box_short";
    tallBox.materialId = "/cornell/materials/box_tall";
    s.meshes.push_back(shortBox);
    s.meshes.push_back(tallBox);
    {
        SplitMix64 r{StreamSeed(seed, 200)};
        s.materials.push_back(Lambert(shortBox.materialId, r));
        SplitMix64 r2{StreamSeed(seed, 201)};
        s.materials.push_back(Lambert(tallBox.materialId, r2));
    }
    SceneLight light;
    light.id = "/cornell/light";
    light.type = SceneLightType::Rect;
    light.color = {1.0F, 1.0F, 1.0F};
    light.intensity = 100.0F;
    light.basisX = {1.0F, 0.0F, 0.0F};
    light.basisY = {0.0F, 0.0F, -1.0F};
    light.basisZ = {0.0F, 1.0F, 0.0F};  // emits along -basisZ = -Y
    light.position = {0.0F, 1.99F, 0.0F};
    light.axisU = {0.25F, 0.0F, 0.0F};
    light.axisV = {0.0F, 0.0F, -0.25F};
    light.width = 0.25F;
    light.height = 0.25F;
    light.area = 0.0625F;
    s.lights.push_back(light);
    SortSnapshot(s);
    out.eye = {0.0, 1.0, 3.4};
    out.target = {0.0, 1.0, 0.0};
    out.up = {0.0, 1.0, 0.0};
    out.verticalFovDegrees = 40.0;
    return out;
}

// ---- 2D gradient noise and fBm for the heightfield ----
struct GradientNoise {
    std::vector<int> perm;
    explicit GradientNoise(std::uint64_t seed) : perm(512)
    {
        std::vector<int> p(256);
        for (int i = 0; i < 256; ++i) p[i] = i;
        SplitMix64 r{seed};
        for (int i = 255; i > 0; --i) {
            const int j = static_cast<int>(r.Next() % static_cast<std::uint64_t>(i + 1));
            std::swap(p[i], p[j]);
        }
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
    }
    static double Fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
    double Grad(int h, double x, double y) const
    {
        // 8 unit-ish gradient directions.
        switch (h & 7) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x;
        case 5: return -x;
        case 6: return y;
        default: return -y;
        }
    }
    double Eval(double x, double y) const
    {
        const double fx = std::floor(x), fy = std::floor(y);
        const int xi = static_cast<int>(fx) & 255, yi = static_cast<int>(fy) & 255;
        const double dx = x - fx, dy = y - fy;
        const double u = Fade(dx), v = Fade(dy);
        const int aa = perm[perm[xi] + yi], ab = perm[perm[xi] + yi + 1];
        const int ba = perm[perm[xi + 1] + yi], bb = perm[perm[xi + 1] + yi + 1];
        const double x1 = Grad(aa, dx, dy) + u * (Grad(ba, dx - 1, dy) - Grad(aa, dx, dy));
        const double x2 = Grad(ab, dx, dy - 1) + u * (Grad(bb, dx - 1, dy - 1) - Grad(ab, dx, dy - 1));
        return x1 + v * (x2 - x1);
    }
    double Fbm(double x, double y, int octaves) const
    {
        double sum = 0.0, amplitude = 0.5, frequency = 1.0;
        for (int o = 0; o < octaves; ++o) {
            sum += amplitude * Eval(x * frequency, y * frequency);
            frequency *= 2.0;  // lacunarity
            amplitude *= 0.5;  // gain
        }
        return sum;
    }
};

SceneTexture SkyDome(std::uint32_t width, std::uint32_t height)
{
    // Lat-long (row 0 = nadir): sky gradient plus a sun disc of peak
    // radiance 5e4 at elevation 35 deg, azimuth 60 deg, radius 0.5 deg with
    // a smooth limb.
    SceneTexture t;
    t.id = "/terrain/sky#hdr";
    t.sourcePath = "procedural";
    t.width = width;
    t.height = height;
    t.rgbaFloat.resize(static_cast<std::size_t>(width) * height * 4);
    const double sunElevation = 35.0 * kPi / 180.0, sunAzimuth = 60.0 * kPi / 180.0;
    const double sun[3] = {std::cos(sunElevation) * std::sin(sunAzimuth), std::sin(sunElevation),
                           -std::cos(sunElevation) * std::cos(sunAzimuth)};
    const double sunRadius = 0.5 * kPi / 180.0;
    for (std::uint32_t y = 0; y < height; ++y) {
        for (std::uint32_t x = 0; x < width; ++x) {
            const double u = (x + 0.5) / width, v = (y + 0.5) / height;
            const double theta = kPi * (1.0 - v), phi = 2.0 * kPi * u;
            const double d[3] = {std::sin(theta) * std::sin(phi), std::cos(theta),
                                 -std::sin(theta) * std::cos(phi)};
            double rgb[3];
            if (d[1] >= 0.0) {
                const double k = std::pow(1.0 - d[1], 3.0);
                rgb[0] = 0.25 + 0.75 * k;
                rgb[1] = 0.45 + 0.55 * k;
                rgb[2] = 0.95 + 0.05 * k;
            } else {
                const double k = std::min(1.0, -d[1] * 4.0);
                rgb[0] = 0.6 * (1 - k) + 0.18 * k;
                rgb[1] = 0.6 * (1 - k) + 0.15 * k;
                rgb[2] = 0.6 * (1 - k) + 0.12 * k;
            }
            const double cosAngle = std::clamp(d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2], -1.0, 1.0);
            const double angle = std::acos(cosAngle);
            if (angle < sunRadius) {
                const double limb = std::sqrt(std::max(0.0, 1.0 - (angle / sunRadius) * (angle / sunRadius)));
                const double peak = 5.0e4 * (0.4 + 0.6 * limb);
                rgb[0] = peak * 1.0;
                rgb[1] = peak * 0.96;
                rgb[2] = peak * 0.90;
            }
            float* out = &t.rgbaFloat[(static_cast<std::size_t>(y) * width + x) * 4];
            out[0] = static_cast<float>(rgb[0]);
            out[1] = static_cast<float>(rgb[1]);
            out[2] = static_cast<float>(rgb[2]);
            out[3] = 1.0F;
        }
    }
    return t;
}

// ---- Heightfield: n x n vertex grid, 2(n-1)^2 triangles, heights from
// seeded 8-octave fBm (lacunarity 2, gain 0.5); 32 x 32-cell material tiles
// of GGX dielectric (roughness U[0.05, 0.6], IOR 1.5) over Burley diffuse;
// lit by an importance-sampled HDR lat-long dome. ----
ProceduralScene MakeHeightfield(std::uint64_t seed, std::uint32_t n, std::uint32_t domeWidth)
{
    if (n < 2) throw std::invalid_argument("heightfield needs at least 2 vertices per side");
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    const GradientNoise noise(StreamSeed(seed, 1));
    const double extent = 100.0, amplitude = 14.0, features = 6.0;
    SceneMesh mesh;
    mesh.id = "/terrain/mesh";
    mesh.positions.resize(static_cast<std::size_t>(n) * n * 3);
    for (std::uint32_t j = 0; j < n; ++j) {
        for (std::uint32_t i = 0; i < n; ++i) {
            const double fx = static_cast<double>(i) / (n - 1), fz = static_cast<double>(j) / (n - 1);
            const double h = amplitude * noise.Fbm(fx * features, fz * features, 8);
            float* p = &mesh.positions[(static_cast<std::size_t>(j) * n + i) * 3];
            p[0] = static_cast<float>((fx - 0.5) * extent);
            p[1] = static_cast<float>(h);
            p[2] = static_cast<float>((fz - 0.5) * extent);
        }
    }
    const std::uint32_t cells = n - 1;
    const std::uint32_t tile = 32;
    const std::uint32_t tilesPerSide = (cells + tile - 1) / tile;
    mesh.indices.reserve(static_cast<std::size_t>(cells) * cells * 6);
    mesh.texcoords.reserve(static_cast<std::size_t>(cells) * cells * 12);
    mesh.triangleMaterialIds.reserve(static_cast<std::size_t>(cells) * cells * 2);
    std::vector<std::string> tileMaterial(static_cast<std::size_t>(tilesPerSide) * tilesPerSide);
    char name[64];
    for (std::uint32_t t = 0; t < tileMaterial.size(); ++t) {
        std::snprintf(name, sizeof(name), "/terrain/materials/tile_%05u", t);
        tileMaterial[t] = name;
        SplitMix64 r{StreamSeed(seed, 1000 + t)};
        SceneMaterial m;
        m.id = name;
        m.baseColor = {r.Range(0.2, 0.8), r.Range(0.2, 0.8), r.Range(0.2, 0.8)};
        const float rough = r.Range(0.05, 0.6);
        m.roughness = rough;
        m.roughnessV = rough;
        m.indexOfRefraction = 1.5F;
        m.specularWeight = 1.0F;
        m.diffuseModel = SceneMaterial::DiffuseModel::Burley;
        m.diffuseRoughness = r.Range(0.0, 1.0);
        s.materials.push_back(m);
    }
    for (std::uint32_t j = 0; j < cells; ++j) {
        for (std::uint32_t i = 0; i < cells; ++i) {
            const std::uint32_t a = j * n + i, b = a + 1, c = a + n, d = c + 1;
            const std::string& mat = tileMaterial[(j / tile) * tilesPerSide + (i / tile)];
            const float u0 = static_cast<float>(i) / cells, u1 = static_cast<float>(i + 1) / cells;
            const float v0 = static_cast<float>(j) / cells, v1 = static_cast<float>(j + 1) / cells;
            mesh.indices.insert(mesh.indices.end(), {a, c, b, b, c, d});
            mesh.texcoords.insert(mesh.texcoords.end(), {u0, v0, u0, v1, u1, v0, u1, v0, u0, v1, u1, v1});
            mesh.triangleMaterialIds.push_back(mat);
            mesh.triangleMaterialIds.push_back(mat);
        }
    }
    mesh.materialId = tileMaterial[0];
    s.meshes.push_back(std::move(mesh));
    const std::uint32_t dw = std::max<std::uint32_t>(domeWidth, 4);
    s.textures.push_back(SkyDome(dw, dw / 2));
    SceneLight dome;
    dome.id = "/terrain/dome";
    dome.type = SceneLightType::Dome;
    dome.texture = "/terrain/sky#hdr";
    dome.textureFormat = DomeTextureFormat::LatLong;
    dome.intensity = 1.0F;
    s.lights.push_back(dome);
    SortSnapshot(s);
    out.eye = {0.0, 25.0, 75.0};
    out.target = {0.0, 8.0, 0.0};
    out.up = {0.0, 1.0, 0.0};
    out.verticalFovDegrees = 42.0;
    return out;
}

// ---- The reference integrator test's one-triangle scene
// (tests/path_tracer_tests.cpp:36-128), camera of :131-136. ----
ProceduralScene MakeTriangle()
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    SceneMesh mesh;
    mesh.id = "/triangle";
    mesh.positions = {-1.25F, -1.0F, -3.0F, 1.25F, -1.0F, -3.0F, 0.0F, 1.2F, -3.0F};
    mesh.indices = {0, 1, 2};
    mesh.texcoords = {0.0F, 0.0F, 1.0F, 0.0F, 0.5F, 1.0F};
    mesh.materialId = "/red";
    s.meshes.push_back(mesh);
    SceneMaterial red;
    red.id = "/red";
    red.baseColor = {1.0F, 1.0F, 1.0F};
    red.metalness = 0.15F;
    red.roughness = 0.3F;
    red.baseColorTexture = "green#srgb";
    s.materials.push_back(red);
    SceneTexture green;
    green.id = "green#srgb";
    green.sourcePath = "synthetic";
    green.width = green.height = 1;
    green.srgb = true;
    green.rgba = {16, 220, 32, 255};
    s.textures.push_back(green);
    SceneTexture scatter;
    scatter.id = "scatter#raw";
    scatter.width = scatter.height = 1;
    scatter.rgba = {255, 255, 255, 255};
    s.textures.push_back(scatter);
    SceneTexture u1;
    u1.id = "synthetic-udim#srgb#udim=1001";
    u1.udimSetId = "synthetic-udim#srgb";
    u1.udimTile = 1001;
    u1.width = u1.height = 1;
    u1.srgb = true;
    u1.rgba = {230, 16, 16, 255};
    s.textures.push_back(u1);
    SceneTexture u2 = u1;
    u2.id = "synthetic-udim#srgb#udim=1002";
    u2.udimTile = 1002;
    u2.rgba = {16, 24, 230, 255};
    s.textures.push_back(u2);
    SceneTexture rect;
    rect.id = "rect#hdr";
    rect.width = rect.height = 1;
    rect.rgbaFloat = {2.0F, 2.0F, 2.0F, 1.0F};
    s.textures.push_back(rect);
    SceneTexture gradient;
    gradient.id = "spectral-gradient#hdr";
    gradient.width = 4;
    gradient.height = 1;
    gradient.rgbaFloat = {3.0F, 0.02F, 0.02F, 1.0F, 0.02F, 3.0F, 0.02F, 1.0F,
                          0.02F, 0.02F, 3.0F, 1.0F, 3.0F, 3.0F, 3.0F, 1.0F};
    s.textures.push_back(gradient);
    SceneLight dome;
    dome.id = "/dome";
    dome.type = SceneLightType::Dome;
    dome.color = {0.08F, 0.16F, 0.8F};
    dome.intensity = 0.35F;
    s.lights.push_back(dome);
    SceneLight r;
    r.id = "/rect";
    r.type = SceneLightType::Rect;
    r.color = {1.0F, 0.92F, 0.82F};
    r.intensity = 50.0F;
    r.texture = "rect#hdr";
    r.basisX = {0.817F, -0.576F, 0.0F};
    r.basisY = {0.184F, 0.261F, -0.947F};
    r.basisZ = {0.547F, 0.775F, 0.319F};
    r.position = {3.282F, 4.650F, -1.086F};
    r.axisU = {1.634F, -1.152F, 0.0F};
    r.axisV = {0.368F, 0.522F, -1.894F};
    r.width = 2.0F;
    r.height = 2.0F;
    r.area = 4.0F;
    s.lights.push_back(r);
    SortSnapshot(s);
    out.explicitCamera = true;
    out.camera.origin = {0.0F, 0.0F, 0.0F};
    out.camera.lowerLeft = {-1.0F, -1.0F, -1.0F};
    out.camera.horizontal = {2.0F, 0.0F, 0.0F};
    out.camera.vertical = {0.0F, 2.0F, 0.0F};
    return out;
}

// ---- Furnace: a 2000 x 2000 Lambert floor (flat albedo 0.5, no specular)
// under a uniform untextured white dome, camera 1 unit above looking down.
// Every camera ray sees the floor; every reflected ray escapes to the dome,
// so the expected pixel value is 0.5 x (dome radiance integrated against
// the sensor) in closed form. ----
ProceduralScene MakeFurnace()
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    SceneMesh floor;
    floor.id = "/furnace/floor";
    const double q[4][3] = {{-1000, 0, 1000}, {1000, 0, 1000}, {1000, 0, -1000}, {-1000, 0, -1000}};
    AddQuad(floor, q);
    floor.materialId = "/furnace/grey";
    s.meshes.push_back(floor);
    SceneMaterial grey;
    grey.id = "/furnace/grey";
    grey.baseColor = {0.5F, 0.5F, 0.5F};
    grey.specularWeight = 0.0F;
    s.materials.push_back(grey);
    SceneLight dome;
    dome.id = "/furnace/dome";
    dome.type = SceneLightType::Dome;
    dome.intensity = 1.0F;
    s.lights.push_back(dome);
    out.eye = {0.0, 1.0, 0.0};
    out.target = {0.0, 0.0, 0.0};
    out.up = {0.0, 0.0, -1.0};
    out.verticalFovDegrees = 60.0;
    return out;
}

// ---- Irradiance cases: one white Lambert receiver (a 40 x 40 quad at
// y = 0, flat albedo 1, no specular weight) and one light whose irradiance
// at the quad's centre has a closed form, seen by a narrow camera 10 units
// above looking straight down (a 1-pixel image covers 2 mm of the quad). Cases (param):
//   0 untextured white dome (intensity 1)          1 rect 1 x 1 at height 2, facing down
//   2 disk of radius 1 at height 3, facing down    3 sphere of radius 0.5 centred at height 3
//   4 distant, delta (angle 0), 20 deg off the normal
//   5 distant, 30 deg cone along the normal        6 case 1 with a 40 deg shaping cone
//   7 rect at height 2 moved 2 units sideways (the centre is 45 deg off its
//     axis) with a 30 deg shaping cone             8 no lights: fallback sun and sky
//   9 the same with a Z-up camera (receiver in the z = 0 plane)
//  10 glass slab 0.5 thick (IOR 1.5, transmission colour 0.6 at depth 1)
//     between y = 0 and y = -0.5 under the white dome
//  11 the slab with a colourless medium (transmission colour 1)
//  12 a closed icosphere of radius 1 at the origin (level 5), white dome ----
ProceduralScene MakeIrradiance(std::uint32_t which)
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    const bool zUp = which == 9;
    SceneMaterial white;
    white.id = "/irradiance/white";
    white.baseColor = {1.0F, 1.0F, 1.0F};
    white.specularWeight = 0.0F;
    if (which >= 10 && which <= 11) {
        SceneMesh slab;
        slab.id = "/irradiance/slab";
        const double top[4][3] = {{-20, 0, 20}, {20, 0, 20}, {20, 0, -20}, {-20, 0, -20}};
        const double bottom[4][3] = {{-20, -0.5, -20}, {20, -0.5, -20}, {20, -0.5, 20}, {-20, -0.5, 20}};
        AddQuad(slab, top);
        AddQuad(slab, bottom);
        slab.materialId = "/irradiance/glass";
        s.meshes.push_back(slab);
        SceneMaterial glass;
        glass.id = "/irradiance/glass";
        glass.baseColor = {1.0F, 1.0F, 1.0F};
        glass.transmission = 1.0F;
        glass.transmissionColor = which == 10 ? std::array<float, 3>{0.6F, 0.6F, 0.6F}
                                              : std::array<float, 3>{1.0F, 1.0F, 1.0F};
        glass.transmissionDepth = 1.0F;
        glass.indexOfRefraction = 1.5F;
        glass.specularWeight = 0.0F;
        glass.roughness = 0.02F;
        glass.roughnessV = 0.02F;
        s.materials.push_back(glass);
    } else if (which == 12) {
        SceneMesh ball;
        ball.id = "/irradiance/ball";
        const double c[3] = {0.0, 0.0, 0.0};
        AppendIcosphere(ball, 5, c, 1.0);
        ball.materialId = "/irradiance/white";
        s.meshes.push_back(ball);
        s.materials.push_back(white);
    } else {
        SceneMesh quad;
        quad.id = "/irradiance/receiver";
        if (zUp) {
            const double q[4][3] = {{-20, -20, 0}, {20, -20, 0}, {20, 20, 0}, {-20, 20, 0}};
            AddQuad(quad, q);
        } else {
            const double q[4][3] = {{-20, 0, 20}, {20, 0, 20}, {20, 0, -20}, {-20, 0, -20}};
            AddQuad(quad, q);
        }
        quad.materialId = "/irradiance/white";
        s.meshes.push_back(quad);
        s.materials.push_back(white);
    }
    SceneLight l;
    l.id = "/irradiance/light";
    l.basisX = {1.0F, 0.0F, 0.0F};
    l.basisY = {0.0F, 0.0F, 1.0F};
    l.basisZ = {0.0F, 1.0F, 0.0F};  // emits along -basisZ: straight down
    switch (which) {
    case 0: case 10: case 11: case 12:
        l.type = SceneLightType::Dome;
        l.basisY = {0.0F, 1.0F, 0.0F};
        l.basisZ = {0.0F, 0.0F, 1.0F};
        break;
    case 1: case 6: case 7:
        l.type = SceneLightType::Rect;
        l.intensity = 5.0F;
        l.position = {which == 7 ? 2.0F : 0.0F, 2.0F, 0.0F};
        l.axisU = {1.0F, 0.0F, 0.0F};
        l.axisV = {0.0F, 0.0F, 1.0F};
        l.width = l.height = 1.0F;
        l.area = 1.0F;
        if (which == 6) l.shapingConeAngle = 40.0F;
        if (which == 7) l.shapingConeAngle = 30.0F;
        break;
    case 2:
        l.type = SceneLightType::Disk;
        l.intensity = 4.0F;
        l.position = {0.0F, 3.0F, 0.0F};
        l.axisU = {1.0F, 0.0F, 0.0F};
        l.axisV = {0.0F, 0.0F, 1.0F};
        l.radius = 1.0F;
        l.area = static_cast<float>(kPi);
        break;
    case 3:
        l.type = SceneLightType::Sphere;
        l.intensity = 6.0F;
        l.position = {0.0F, 3.0F, 0.0F};
        l.radius = 0.5F;
        l.area = static_cast<float>(4.0 * kPi * 0.25);
        break;
    case 4: {
        l.type = SceneLightType::Distant;
        l.intensity = 2.0F;
        l.angle = 0.0F;
        const double a = 20.0 * kPi / 180.0;  // direction toward the light: 20 deg from +y in the xy plane
        l.basisZ = {static_cast<float>(std::sin(a)), static_cast<float>(std::cos(a)), 0.0F};
        l.basisX = {static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a)), 0.0F};
        break;
    }
    case 5:
        l.type = SceneLightType::Distant;
        l.intensity = 2.0F;
        l.angle = 60.0F;  // full cone angle: 30 deg half angle
        l.basisX = {1.0F, 0.0F, 0.0F};
        l.basisY = {0.0F, 0.0F, -1.0F};
        l.basisZ = {0.0F, 1.0F, 0.0F};
        break;
    default:
        break;
    }
    if (which != 8 && which != 9) s.lights.push_back(l);
    SortSnapshot(s);
    out.explicitCamera = true;
    if (zUp) {
        // Looking straight down -z; the vertical vector has the larger z
        // component, which makes the camera Z-up for the fallback sky
        // (vulkan_path_tracer.cpp:3237-3241).
        out.camera.origin = {0.0F, 0.0F, 10.0F};
        out.camera.lowerLeft = {-0.0001F, -0.0001F, 9.0F - 0.000105F};
        out.camera.horizontal = {0.0002F, 0.0F, 0.0F};
        out.camera.vertical = {0.0F, 0.0002F, 0.00021F};
    } else {
        // Looking straight down -y; the vertical vector has the larger y
        // component, so the camera is Y-up for the fallback sky.
        out.camera.origin = {0.0F, 10.0F, 0.0F};
        out.camera.lowerLeft = {-0.0001F, 9.0F - 0.000105F, 0.0001F};
        out.camera.horizontal = {0.0002F, 0.0F, 0.0F};
        out.camera.vertical = {0.0F, 0.00021F, -0.0002F};
    }
    return out;
}

} // namespace

PathTracerCamera ProceduralScene::Camera(std::uint32_t width, std::uint32_t height) const
{
    if (explicitCamera) return camera;
    // Pinhole look-at camera at unit focal distance.
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& c : f) c /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& c : r) c /= rl;
    const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    const double vh = 2.0 * std::tan(verticalFovDegrees * kPi / 360.0);
    const double hw = vh * static_cast<double>(width) / static_cast<double>(std::max(height, 1U));
    PathTracerCamera c;
    for (int k = 0; k < 3; ++k) {
        c.origin[k] = static_cast<float>(eye[k]);
        c.horizontal[k] = static_cast<float>(r[k] * hw);
        c.vertical[k] = static_cast<float>(u[k] * vh);
        c.lowerLeft[k] = static_cast<float>(eye[k] + f[k] - 0.5 * r[k] * hw - 0.5 * u[k] * vh);
    }
    return c;
}

ProceduralScene MakeProceduralScene(const std::string& name, std::uint64_t seed,
                                    std::uint32_t param, std::uint32_t domeWidth)
{
    if (name == "cornell") return MakeCornell(seed);
    if (name == "heightfield") return MakeHeightfield(seed, param == 0 ? 1024 : param,
                                                      domeWidth == 0 ? 2048 : domeWidth);
    if (name == "triangle") return MakeTriangle();
    if (name == "materials") return MakeMaterialShowcase(seed);
    if (name == "furnace") return MakeFurnace();
    if (name == "irradiance") return MakeIrradiance(param);
    if (name == "mixed") return MakeMixedScene(seed, param == 0 ? 20000 : param, domeWidth == 0 ? 256 : domeWidth, 8);
    if (name == "instanced") return MakeInstancedScene(seed, param == 0 ? 10000 : param, 5, domeWidth == 0 ? 256 : domeWidth, 64);
    if (name == "instanced_small") return MakeInstancedScene(seed, param == 0 ? 50 : param, 3, domeWidth == 0 ? 64 : domeWidth, 8);
    if (name == "mixed_lights") return MakeMixedScene(seed, param == 0 ? 20000 : param, domeWidth == 0 ? 256 : domeWidth, 63);
    throw std::invalid_argument("unknown procedural scene '" + name + "'");
}

} // namespace hdcx
