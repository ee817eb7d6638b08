// "materials" scene: one small opaque scene that reaches every closure lobe,
// diffuse model, light type, shaping/shadow control and texture path of the
// integrator, so that CPU/GPU parity covers the whole shading code.
#include "procedural.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <utility>

namespace hdcx {
namespace {

constexpr double kPi = 3.14159265358979323846;

SceneMesh Icosphere(const std::string& id, int level, const double c[3], double radius)
{
    SceneMesh m;
    m.id = id;
    AppendIcosphere(m, level, c, radius);
    return m;
}

SceneMesh Quad(const std::string& id, const double p[4][3], float uvScale)
{
    SceneMesh m;
    m.id = id;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) m.positions.push_back(static_cast<float>(p[i][k]));
    m.indices = {0, 1, 2, 0, 2, 3};
    const float uv[4][2] = {{0, 0}, {uvScale, 0}, {uvScale, uvScale}, {0, uvScale}};
    const int corners[6] = {0, 1, 2, 0, 2, 3};
    for (int c : corners) m.texcoords.insert(m.texcoords.end(), {uv[c][0], uv[c][1]});
    return m;
}

SceneTexture Checker(const std::string& id, std::uint32_t size, bool srgb,
                     std::array<std::uint8_t, 3> a, std::array<std::uint8_t, 3> b)
{
    SceneTexture t;
    t.id = id;
    t.sourcePath = "procedural";
    t.width = t.height = size;
    t.srgb = srgb;
    t.rgba.resize(static_cast<std::size_t>(size) * size * 4);
    for (std::uint32_t y = 0; y < size; ++y)
        for (std::uint32_t x = 0; x < size; ++x) {
            const bool odd = ((x / 4) + (y / 4)) & 1U;
            const auto& c = odd ? a : b;
            std::uint8_t* o = &t.rgba[(static_cast<std::size_t>(y) * size + x) * 4];
            o[0] = c[0];
            o[1] = c[1];
            o[2] = c[2];
            o[3] = 255;
        }
    return t;
}

} // namespace

ProceduralScene MakeMaterialShowcase(std::uint64_t seed)
{
    (void)seed;
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;

    // Textures.
    s.textures.push_back(Checker("/tex/checker#srgb", 32, true, {200, 190, 170}, {60, 70, 90}));
    {
        // Tangent-space normal map: bumps in x.
        SceneTexture n;
        n.id = "/tex/bumps#raw";
        n.width = 16;
        n.height = 16;
        n.rgba.resize(16 * 16 * 4);
        for (std::uint32_t y = 0; y < 16; ++y)
            for (std::uint32_t x = 0; x < 16; ++x) {
                const double sx = std::sin(2.0 * kPi * x / 16.0) * 0.5;
                const double sy = std::cos(2.0 * kPi * y / 16.0) * 0.3;
                const double nz = std::sqrt(std::max(0.0, 1.0 - sx * sx - sy * sy));
                std::uint8_t* o = &n.rgba[(y * 16 + x) * 4];
                o[0] = static_cast<std::uint8_t>(std::lround((sx * 0.5 + 0.5) * 255));
                o[1] = static_cast<std::uint8_t>(std::lround((sy * 0.5 + 0.5) * 255));
                o[2] = static_cast<std::uint8_t>(std::lround((nz * 0.5 + 0.5) * 255));
                o[3] = 255;
            }
        s.textures.push_back(n);
    }
    {
        SceneTexture u1 = Checker("/tex/udim#srgb#udim=1001", 8, true, {230, 40, 30}, {240, 220, 200});
        u1.udimSetId = "/tex/udim#srgb";
        u1.udimTile = 1001;
        SceneTexture u2 = Checker("/tex/udim#srgb#udim=1002", 8, true, {30, 60, 230}, {200, 220, 240});
        u2.udimSetId = "/tex/udim#srgb";
        u2.udimTile = 1002;
        s.textures.push_back(u1);
        s.textures.push_back(u2);
    }
    {
        SceneTexture sky;
        sky.id = "/tex/sky#hdr";
        sky.width = 64;
        sky.height = 32;
        sky.rgbaFloat.resize(64 * 32 * 4);
        for (std::uint32_t y = 0; y < 32; ++y)
            for (std::uint32_t x = 0; x < 64; ++x) {
                float* o = &sky.rgbaFloat[(y * 64 + x) * 4];
                const double v = (y + 0.5) / 32.0;
                o[0] = static_cast<float>(0.2 + 0.5 * v);
                o[1] = static_cast<float>(0.25 + 0.45 * v);
                o[2] = static_cast<float>(0.4 + 0.6 * v);
                o[3] = 1.0F;
                if (x >= 40 && x < 43 && y >= 24 && y < 26) {
                    o[0] = 60.0F;
                    o[1] = 55.0F;
                    o[2] = 45.0F;
                }
            }
        s.textures.push_back(sky);
    }
    {
        SceneTexture rt;
        rt.id = "/tex/rect#hdr";
        rt.width = 2;
        rt.height = 2;
        rt.rgbaFloat = {2.0F, 1.5F, 1.0F, 1.0F, 1.0F, 1.5F, 2.0F, 1.0F,
                        1.5F, 1.5F, 1.5F, 1.0F, 2.0F, 2.0F, 2.0F, 1.0F};
        s.textures.push_back(rt);
    }

    // Floor: checker base color, Oren-Nayar qualitative.
    const double floorQuad[4][3] = {{-6, 0, 6}, {6, 0, 6}, {6, 0, -6}, {-6, 0, -6}};
    SceneMesh floor = Quad("/geo/floor", floorQuad, 4.0F);
    floor.materialId = "/mat/floor";
    s.meshes.push_back(floor);
    SceneMaterial floorMat;
    floorMat.id = "/mat/floor";
    floorMat.baseColorTexture = "/tex/checker#srgb";
    floorMat.diffuseModel = SceneMaterial::DiffuseModel::OrenNayar;
    floorMat.diffuseRoughness = 0.6F;
    floorMat.specularWeight = 0.3F;
    floorMat.roughness = 0.45F;
    floorMat.roughnessV = 0.45F;
    s.materials.push_back(floorMat);

    // Back wall: UDIM texture + normal map, EON diffuse.
    const double wallQuad[4][3] = {{-6, 0, -3}, {6, 0, -3}, {6, 5, -3}, {-6, 5, -3}};
    SceneMesh wall = Quad("/geo/wall", wallQuad, 2.0F);
    wall.materialId = "/mat/wall";
    s.meshes.push_back(wall);
    SceneMaterial wallMat;
    wallMat.id = "/mat/wall";
    wallMat.baseColorTexture = "/tex/udim#srgb";
    wallMat.normalTexture = "/tex/bumps#raw";
    wallMat.normalTextureScale = 0.8F;
    wallMat.normalTextureFlipY = true;
    wallMat.diffuseModel = SceneMaterial::DiffuseModel::OrenNayarEnergyCompensated;
    wallMat.diffuseRoughness = 0.8F;
    wallMat.specularWeight = 0.5F;
    s.materials.push_back(wallMat);

    struct Ball {
        const char* name;
        double c[3];
    };
    const Ball balls[9] = {
        {"a_conductor", {-3.6, 0.6, 0.0}}, {"b_glass", {-2.4, 0.6, 0.0}}, {"c_subsurface", {-1.2, 0.6, 0.0}},
        {"d_translucent", {0.0, 0.6, 0.0}}, {"e_sheen", {1.2, 0.6, 0.0}}, {"f_coat", {2.4, 0.6, 0.0}},
        {"g_thin", {3.6, 0.6, 0.0}}, {"h_emitter", {-1.8, 1.9, -1.2}}, {"i_burley", {1.8, 1.9, -1.2}}};
    for (const Ball& b : balls) {
        SceneMesh m = Icosphere(std::string("/geo/ball_") + b.name, 2, b.c, 0.5);
        m.materialId = std::string("/mat/ball_") + b.name;
        s.meshes.push_back(m);
    }
    SceneMaterial conductor;
    conductor.id = "/mat/ball_a_conductor";
    conductor.baseColor = {0.95F, 0.64F, 0.54F};
    conductor.metalness = 1.0F;
    conductor.roughness = 0.2F;
    conductor.roughnessV = 0.45F;
    s.materials.push_back(conductor);
    SceneMaterial glass;
    glass.id = "/mat/ball_b_glass";
    glass.baseColor = {1.0F, 1.0F, 1.0F};
    glass.transmission = 1.0F;
    glass.transmissionColor = {0.7F, 0.9F, 0.8F};
    glass.transmissionDepth = 0.5F;
    glass.transmissionScatter = {0.3F, 0.3F, 0.3F};
    glass.transmissionScatterAnisotropy = 0.3F;
    glass.transmissionDispersionScale = 1.0F;
    glass.transmissionDispersionAbbeNumber = 30.0F;
    glass.indexOfRefraction = 1.5F;
    glass.roughness = 0.05F;
    glass.roughnessV = 0.05F;
    s.materials.push_back(glass);
    SceneMaterial sss;
    sss.id = "/mat/ball_c_subsurface";
    sss.baseColor = {0.9F, 0.6F, 0.5F};
    sss.subsurface = 1.0F;
    sss.subsurfaceColor = {0.9F, 0.5F, 0.4F};
    sss.subsurfaceRadius = {0.4F, 0.2F, 0.1F};
    sss.subsurfaceScale = 0.5F;
    sss.subsurfaceScatterAnisotropy = 0.2F;
    sss.roughness = 0.4F;
    sss.roughnessV = 0.4F;
    s.materials.push_back(sss);
    SceneMaterial translucent;
    translucent.id = "/mat/ball_d_translucent";
    translucent.baseColor = {0.3F, 0.8F, 0.3F};
    translucent.translucentWeight = 0.7F;
    translucent.translucentColor = {0.4F, 0.9F, 0.3F};
    translucent.roughness = 0.5F;
    translucent.roughnessV = 0.5F;
    s.materials.push_back(translucent);
    SceneMaterial sheen;
    sheen.id = "/mat/ball_e_sheen";
    sheen.baseColor = {0.5F, 0.2F, 0.6F};
    sheen.sheen = 1.0F;
    sheen.sheenColor = {0.9F, 0.8F, 1.0F};
    sheen.sheenRoughness = 0.35F;
    sheen.sheenMode = 1U;
    sheen.diffuseModel = SceneMaterial::DiffuseModel::OrenNayarEnergyCompensated;
    sheen.diffuseRoughness = 0.5F;
    s.materials.push_back(sheen);
    SceneMaterial coat;
    coat.id = "/mat/ball_f_coat";
    coat.baseColor = {0.7F, 0.1F, 0.1F};
    coat.coat = 1.0F;
    coat.coatRoughness = 0.05F;
    coat.coatRoughnessV = 0.05F;
    coat.coatColor = {1.0F, 0.95F, 0.9F};
    coat.coatIndexOfRefraction = 1.6F;
    coat.sheen = 0.3F;
    coat.sheenMode = 0U;
    coat.roughness = 0.6F;
    coat.roughnessV = 0.6F;
    s.materials.push_back(coat);
    SceneMaterial thin;
    thin.id = "/mat/ball_g_thin";
    thin.transmission = 0.8F;
    thin.thinWalled = true;
    thin.transmissionColor = {0.9F, 0.7F, 0.3F};
    thin.roughness = 0.1F;
    thin.roughnessV = 0.1F;
    s.materials.push_back(thin);
    SceneMaterial emitter;
    emitter.id = "/mat/ball_h_emitter";
    emitter.baseColor = {0.2F, 0.2F, 0.2F};
    emitter.emission = {1.0F, 0.8F, 0.6F};
    emitter.emissionWeight = 6.0F;
    s.materials.push_back(emitter);
    SceneMaterial burley;
    burley.id = "/mat/ball_i_burley";
    burley.baseColor = {0.8F, 0.7F, 0.2F};
    burley.diffuseModel = SceneMaterial::DiffuseModel::Burley;
    burley.diffuseRoughness = 0.7F;
    burley.metalness = 0.3F;
    burley.specularColor = {1.0F, 0.9F, 0.8F};
    burley.specularWeight = 0.8F;
    burley.roughness = 0.3F;
    burley.roughnessV = 0.3F;
    s.materials.push_back(burley);

    // Lights: every type, with shaping and shadow controls.
    SceneLight dome;
    dome.id = "/light/a_dome";
    dome.type = SceneLightType::Dome;
    dome.texture = "/tex/sky#hdr";
    dome.intensity = 0.6F;
    s.lights.push_back(dome);
    SceneLight rect;
    rect.id = "/light/b_rect";
    rect.type = SceneLightType::Rect;
    rect.texture = "/tex/rect#hdr";
    rect.intensity = 8.0F;
    rect.basisX = {1.0F, 0.0F, 0.0F};
    rect.basisY = {0.0F, 0.0F, -1.0F};
    rect.basisZ = {0.0F, 1.0F, 0.0F};
    rect.position = {0.0F, 4.0F, 1.0F};
    rect.axisU = {1.5F, 0.0F, 0.0F};
    rect.axisV = {0.0F, 0.0F, -1.0F};
    rect.area = 1.5F;
    rect.shapingFocus = 2.0F;
    rect.shapingFocusTint = {0.2F, 0.3F, 1.0F};
    rect.shapingConeAngle = 70.0F;
    rect.shapingConeSoftness = 0.4F;
    s.lights.push_back(rect);
    SceneLight disk;
    disk.id = "/light/c_disk";
    disk.type = SceneLightType::Disk;
    disk.intensity = 12.0F;
    disk.color = {1.0F, 0.6F, 0.3F};
    disk.basisX = {1.0F, 0.0F, 0.0F};
    disk.basisY = {0.0F, 0.70710678F, -0.70710678F};
    disk.basisZ = {0.0F, 0.70710678F, 0.70710678F};
    disk.position = {-3.0F, 3.0F, 3.0F};
    disk.axisU = {0.4F, 0.0F, 0.0F};
    disk.axisV = {0.0F, 0.28284271F, -0.28284271F};
    disk.radius = 0.4F;
    disk.area = static_cast<float>(kPi * 0.16);
    disk.shadowColor = {0.3F, 0.0F, 0.0F};
    s.lights.push_back(disk);
    SceneLight sphere;
    sphere.id = "/light/d_sphere";
    sphere.type = SceneLightType::Sphere;
    sphere.intensity = 20.0F;
    sphere.normalize = true;
    sphere.color = {0.6F, 0.8F, 1.0F};
    sphere.position = {3.0F, 2.5F, 2.0F};
    sphere.radius = 0.25F;
    sphere.area = static_cast<float>(4.0 * kPi * 0.0625);
    sphere.shadowDistance = 3.0F;
    sphere.shadowFalloff = 1.0F;
    sphere.shadowFalloffGamma = 2.0F;
    s.lights.push_back(sphere);
    SceneLight cylinder;
    cylinder.id = "/light/e_cylinder";
    cylinder.type = SceneLightType::Cylinder;
    cylinder.intensity = 6.0F;
    cylinder.position = {0.0F, 3.2F, -2.0F};
    cylinder.basisX = {1.0F, 0.0F, 0.0F};
    cylinder.basisY = {0.0F, 1.0F, 0.0F};
    cylinder.basisZ = {0.0F, 0.0F, 1.0F};
    cylinder.axisU = {2.0F, 0.0F, 0.0F};
    cylinder.radius = 0.08F;
    cylinder.length = 2.0F;
    cylinder.area = static_cast<float>(2.0 * kPi * 0.08 * 2.0);
    s.lights.push_back(cylinder);
    SceneLight distant;
    distant.id = "/light/f_distant";
    distant.type = SceneLightType::Distant;
    distant.intensity = 1.5F;
    distant.color = {1.0F, 0.95F, 0.85F};
    distant.basisX = {1.0F, 0.0F, 0.0F};
    distant.basisY = {0.0F, 0.5F, -0.8660254F};
    distant.basisZ = {0.0F, 0.8660254F, 0.5F};
    distant.angle = 3.0F;
    distant.diffuse = 0.8F;
    distant.specular = 0.5F;
    s.lights.push_back(distant);

    std::sort(s.meshes.begin(), s.meshes.end(), [](const SceneMesh& a, const SceneMesh& b) { return a.id < b.id; });
    std::sort(s.materials.begin(), s.materials.end(),
              [](const SceneMaterial& a, const SceneMaterial& b) { return a.id < b.id; });
    std::sort(s.textures.begin(), s.textures.end(),
              [](const SceneTexture& a, const SceneTexture& b) { return a.id < b.id; });
    std::sort(s.lights.begin(), s.lights.end(), [](const SceneLight& a, const SceneLight& b) { return a.id < b.id; });
    out.eye = {0.0, 2.6, 8.5};
    out.target = {0.0, 0.8, 0.0};
    out.up = {0.0, 1.0, 0.0};
    out.verticalFovDegrees = 38.0;
    return out;
}

} // namespace hdcx
