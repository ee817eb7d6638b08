// "mixed" scenes: icospheres on a ground plane with the six-material mix of
// the mixed-material and batched workloads, rect/sphere/disk lights and a
// sky dome. Also the shared unit icosphere.
#include "procedural.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <utility>

namespace hdcx {

namespace {

constexpr double kPi = 3.14159265358979323846;

IcoMesh BuildUnitIcosphere(int level)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    IcoMesh m;
    m.vertices = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                  {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : m.vertices) {
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& x : p) x /= l;
    }
    m.faces = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
               {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
               {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<std::uint32_t, std::uint32_t>, std::uint32_t> mid;
        auto midpoint = [&](std::uint32_t a, std::uint32_t b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            if (auto it = mid.find(key); it != mid.end()) return it->second;
            const auto& va = m.vertices[a];
            const auto& vb = m.vertices[b];
            std::array<double, 3> p{(va[0] + vb[0]) / 2, (va[1] + vb[1]) / 2, (va[2] + vb[2]) / 2};
            const double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (double& x : p) x /= len;
            m.vertices.push_back(p);
            const auto index = static_cast<std::uint32_t>(m.vertices.size() - 1);
            mid.emplace(key, index);
            return index;
        };
        std::vector<std::array<std::uint32_t, 3>> next;
        next.reserve(m.faces.size() * 4);
        for (const auto& tri : m.faces) {
            const std::uint32_t a = midpoint(tri[0], tri[1]);
            const std::uint32_t b = midpoint(tri[1], tri[2]);
            const std::uint32_t c = midpoint(tri[2], tri[0]);
            next.push_back({tri[0], a, c});
            next.push_back({tri[1], b, a});
            next.push_back({tri[2], c, b});
            next.push_back({a, b, c});
        }
        m.faces = std::move(next);
    }
    return m;
}

void Sky(SceneTexture& t, std::uint32_t w, std::uint32_t h, SplitMix64& rng)
{
    t.width = w;
    t.height = h;
    t.rgbaFloat.resize(static_cast<std::size_t>(w) * h * 4);
    const double sunU = rng.Uniform(), sunV = 0.62 + 0.25 * rng.Uniform();
    for (std::uint32_t y = 0; y < h; ++y)
        for (std::uint32_t x = 0; x < w; ++x) {
            const double v = (y + 0.5) / h, u = (x + 0.5) / w;
            float* o = &t.rgbaFloat[(static_cast<std::size_t>(y) * w + x) * 4];
            const double up = std::max(0.0, v - 0.5) * 2.0;
            o[0] = static_cast<float>(v < 0.5 ? 0.25 : 0.5 + 0.3 * (1.0 - up));
            o[1] = static_cast<float>(v < 0.5 ? 0.22 : 0.6 + 0.25 * (1.0 - up));
            o[2] = static_cast<float>(v < 0.5 ? 0.2 : 0.9 + 0.1 * up);
            o[3] = 1.0F;
            double du = std::fabs(u - sunU);
            du = std::min(du, 1.0 - du) * 2.0;
            const double dv = v - sunV;
            if (du * du + dv * dv < 0.0004) {
                o[0] = 400.0F;
                o[1] = 380.0F;
                o[2] = 340.0F;
            }
        }
}

SceneMaterial MixMaterial(const std::string& id, SplitMix64& rng)
{
    SceneMaterial mat;
    mat.id = id;
    mat.baseColor = {rng.Range(0.1, 0.9), rng.Range(0.1, 0.9), rng.Range(0.1, 0.9)};
    switch (rng.Next() % 6U) {
    case 0:  // Lambert
        mat.specularWeight = 0.0F;
        break;
    case 1:  // Oren-Nayar
        mat.diffuseModel = SceneMaterial::DiffuseModel::OrenNayar;
        mat.diffuseRoughness = rng.Range(0.3, 1.0);
        mat.specularWeight = 0.0F;
        break;
    case 2:  // GGX conductor
        mat.metalness = 1.0F;
        mat.baseColor = {rng.Range(0.5, 1.0), rng.Range(0.5, 1.0), rng.Range(0.5, 1.0)};
        mat.roughness = rng.Range(0.05, 0.5);
        mat.roughnessV = mat.roughness;
        break;
    case 3:  // dispersive dielectric
        mat.baseColor = {1.0F, 1.0F, 1.0F};
        mat.transmission = 1.0F;
        mat.transmissionColor = {1.0F, 1.0F, 1.0F};
        mat.transmissionDispersionScale = 1.0F;
        mat.transmissionDispersionAbbeNumber = rng.Range(20.0, 60.0);
        mat.indexOfRefraction = 1.5F;
        mat.roughness = rng.Range(0.02, 0.1);
        mat.roughnessV = mat.roughness;
        break;
    case 4:  // diffuse translucent
        mat.translucentWeight = 1.0F;
        mat.translucentColor = {rng.Range(0.3, 0.9), rng.Range(0.3, 0.9), rng.Range(0.3, 0.9)};
        mat.specularWeight = 0.2F;
        break;
    default:  // subsurface random walk
        mat.subsurface = 1.0F;
        mat.subsurfaceColor = {rng.Range(0.4, 0.95), rng.Range(0.4, 0.95), rng.Range(0.4, 0.95)};
        mat.subsurfaceRadius = {rng.Range(0.01, 0.5), rng.Range(0.01, 0.5), rng.Range(0.01, 0.5)};
        mat.subsurfaceScale = 1.0F;
        mat.roughness = 0.4F;
        mat.roughnessV = 0.4F;
        break;
    }
    return mat;
}

void AddSkyAndLights(SceneSnapshot& s, const std::string& prefix, std::uint32_t domeWidth, std::uint32_t lights,
                     double extent, double spacing, SplitMix64& rng)
{
    if (domeWidth > 0) {
        SceneTexture sky;
        sky.id = prefix + "/tex/sky";
        Sky(sky, domeWidth, std::max(domeWidth / 2, 1U), rng);
        s.textures.push_back(sky);
        SceneLight dome;
        dome.id = prefix + "/light/a_dome";
        dome.type = SceneLightType::Dome;
        dome.texture = prefix + "/tex/sky";
        dome.intensity = 0.5F;
        s.lights.push_back(dome);
    }
    for (std::uint32_t i = 0; i < lights; ++i) {
        char name[48];
        std::snprintf(name, sizeof(name), "/light/b_%04u", i);
        SceneLight l;
        l.id = prefix + name;
        l.color = {rng.Range(0.6, 1.0), rng.Range(0.6, 1.0), rng.Range(0.6, 1.0)};
        const float x = static_cast<float>((rng.Uniform() - 0.5) * extent);
        const float z = static_cast<float>((rng.Uniform() - 0.5) * extent);
        const float y = static_cast<float>(spacing * (2.0 + 2.0 * rng.Uniform()));
        l.position = {x, y, z};
        l.basisX = {1.0F, 0.0F, 0.0F};
        l.basisY = {0.0F, 0.0F, -1.0F};
        l.basisZ = {0.0F, 1.0F, 0.0F};  // emits along -basisZ (down)
        switch (i % 3U) {
        case 0: {
            l.type = SceneLightType::Rect;
            const float a = rng.Range(0.3, 1.0), b = rng.Range(0.3, 1.0);
            l.axisU = {a, 0.0F, 0.0F};
            l.axisV = {0.0F, 0.0F, -b};
            l.area = a * b;
            l.intensity = rng.Range(5.0, 20.0);
            break;
        }
        case 1:
            l.type = SceneLightType::Sphere;
            l.radius = rng.Range(0.1, 0.4);
            l.area = static_cast<float>(4.0 * kPi * l.radius * l.radius);
            l.intensity = rng.Range(5.0, 20.0);
            break;
        default: {
            l.type = SceneLightType::Disk;
            l.radius = rng.Range(0.2, 0.6);
            l.axisU = {l.radius, 0.0F, 0.0F};
            l.axisV = {0.0F, 0.0F, -l.radius};
            l.area = static_cast<float>(kPi * l.radius * l.radius);
            l.intensity = rng.Range(5.0, 20.0);
            break;
        }
        }
        s.lights.push_back(l);
    }
}

} // namespace

const IcoMesh& UnitIcosphere(int level)
{
    static std::mutex mu;
    static std::map<int, IcoMesh> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(level);
    if (it == cache.end()) it = cache.emplace(level, BuildUnitIcosphere(level)).first;
    return it->second;
}

void AppendIcosphere(SceneMesh& m, int level, const double c[3], double radius)
{
    const IcoMesh& u = UnitIcosphere(level);
    const auto base = static_cast<std::uint32_t>(m.positions.size() / 3);
    for (const auto& p : u.vertices)
        for (int k = 0; k < 3; ++k) m.positions.push_back(static_cast<float>(c[k] + radius * p[k]));
    for (const auto& tri : u.faces)
        for (int k = 0; k < 3; ++k) {
            const auto& p = u.vertices[tri[k]];
            m.indices.push_back(base + tri[k]);
            m.normals.insert(m.normals.end(), {static_cast<float>(p[0]), static_cast<float>(p[1]),
                                               static_cast<float>(p[2])});
            const double uu = 0.5 + std::atan2(p[0], p[2]) / (2.0 * kPi);
            const double vv = 0.5 + std::asin(std::max(-1.0, std::min(1.0, p[1]))) / kPi;
            m.texcoords.insert(m.texcoords.end(), {static_cast<float>(2.0 * uu), static_cast<float>(vv)});
        }
}

ProceduralScene MakeMixedScene(std::uint64_t seed, std::uint32_t triangleTarget, std::uint32_t domeWidth,
                               std::uint32_t lights)
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    SplitMix64 rng{StreamSeed(seed, 300)};

    // Sphere level: the finest whose triangle count (20 * 4^l) leaves at
    // least four spheres, then as many spheres as the target asks for.
    const std::uint32_t target = std::max<std::uint32_t>(triangleTarget, 100U);
    int level = 1;
    while (level < 7 && 20U * (1U << (2 * (level + 1))) * 4U <= target) ++level;
    const std::uint32_t perSphere = 20U * (1U << (2 * level));
    const std::uint32_t spheres = std::max<std::uint32_t>(1U, (target + perSphere / 2) / perSphere);
    const auto side = static_cast<std::uint32_t>(std::ceil(std::sqrt(static_cast<double>(spheres))));
    const double spacing = 2.0, extent = spacing * side;

    {
        SceneMesh ground;
        ground.id = "/mixed/a_ground";
        const double g = extent * 2.0;
        const float q[4][3] = {{static_cast<float>(-g), 0, static_cast<float>(g)},
                               {static_cast<float>(g), 0, static_cast<float>(g)},
                               {static_cast<float>(g), 0, static_cast<float>(-g)},
                               {static_cast<float>(-g), 0, static_cast<float>(-g)}};
        for (const auto& p : q) ground.positions.insert(ground.positions.end(), {p[0], p[1], p[2]});
        ground.indices = {0, 1, 2, 0, 2, 3};
        ground.materialId = "/mixed/mat/a_ground";
        s.meshes.push_back(ground);
        SceneMaterial gm;
        gm.id = "/mixed/mat/a_ground";
        gm.baseColor = {0.45F, 0.42F, 0.4F};
        gm.specularWeight = 0.0F;
        s.materials.push_back(gm);
    }
    for (std::uint32_t i = 0; i < spheres; ++i) {
        char name[48];
        std::snprintf(name, sizeof(name), "b_sphere%06u", i);
        const double r = spacing * (0.25 + 0.2 * rng.Uniform());
        const double cx = (static_cast<double>(i % side) + 0.5) * spacing - 0.5 * extent + 0.3 * (rng.Uniform() - 0.5);
        const double cz = (static_cast<double>(i / side) + 0.5) * spacing - 0.5 * extent + 0.3 * (rng.Uniform() - 0.5);
        const double c[3] = {cx, r, cz};
        SceneMesh m;
        m.id = std::string("/mixed/") + name;
        AppendIcosphere(m, level, c, r);
        m.materialId = std::string("/mixed/mat/") + name;
        s.meshes.push_back(std::move(m));

        s.materials.push_back(MixMaterial(std::string("/mixed/mat/") + name, rng));
    }

    AddSkyAndLights(s, "/mixed", domeWidth, lights, extent, spacing, rng);

    std::sort(s.meshes.begin(), s.meshes.end(), [](const SceneMesh& a, const SceneMesh& b) { return a.id < b.id; });
    std::sort(s.materials.begin(), s.materials.end(),
              [](const SceneMaterial& a, const SceneMaterial& b) { return a.id < b.id; });
    std::sort(s.lights.begin(), s.lights.end(), [](const SceneLight& a, const SceneLight& b) { return a.id < b.id; });
    out.eye = {0.0, extent * 0.55, extent * 0.95};
    out.target = {0.0, spacing * 0.3, 0.0};
    out.up = {0.0, 1.0, 0.0};
    out.verticalFovDegrees = 45.0;
    return out;
}

ProceduralScene MakeInstancedScene(std::uint64_t seed, std::uint32_t instances, int level, std::uint32_t domeWidth,
                                   std::uint32_t lights)
{
    ProceduralScene out;
    SceneSnapshot& s = out.snapshot;
    s.revision = 1;
    SplitMix64 rng{StreamSeed(seed, 400)};
    const std::uint32_t n = std::max<std::uint32_t>(instances, 1U);
    const auto side = static_cast<std::uint32_t>(std::ceil(std::cbrt(static_cast<double>(n) * 4.0)));
    const double spacing = 2.0, extent = spacing * side;
    {
        SceneMesh ground;
        ground.id = "/inst/a_ground";
        const auto g = static_cast<float>(extent * 2.0);
        ground.positions = {-g, 0, g, g, 0, g, g, 0, -g, -g, 0, -g};
        ground.indices = {0, 1, 2, 0, 2, 3};
        ground.materialId = "/inst/mat/a_ground";
        s.meshes.push_back(ground);
        SceneMaterial gm;
        gm.id = "/inst/mat/a_ground";
        gm.baseColor = {0.45F, 0.42F, 0.4F};
        gm.specularWeight = 0.0F;
        s.materials.push_back(gm);
    }
    {
        SceneMesh proto;
        proto.id = "/inst/b_sphere";
        const double c[3] = {0.0, 0.0, 0.0};
        AppendIcosphere(proto, level, c, 1.0);
        proto.materialId = "/inst/mat/a_ground";
        proto.prototype = true;
        s.meshes.push_back(std::move(proto));
    }
    // Random TRS: uniform scale, uniform rotation (a random unit
    // quaternion), positions in a slab above the ground.
    for (std::uint32_t i = 0; i < n; ++i) {
        char name[48];
        std::snprintf(name, sizeof(name), "%07u", i);
        const double scale = spacing * (0.15 + 0.25 * rng.Uniform());
        double q[4] = {0, 0, 0, 0}, len = 0.0;
        do {
            for (double& x : q) x = 2.0 * rng.Uniform() - 1.0;
            len = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        } while (len > 1.0 || len < 1e-6);
        for (double& x : q) x /= std::sqrt(len);
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                             2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                             2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
        const double t[3] = {(rng.Uniform() - 0.5) * extent, scale + rng.Uniform() * extent * 0.25,
                             (rng.Uniform() - 0.5) * extent};
        SceneInstance inst;
        inst.id = std::string("/inst/i_") + name;
        inst.prototypeMeshId = "/inst/b_sphere";
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col) inst.objectToWorld[4 * row + col] = static_cast<float>(scale * r[3 * row + col]);
            inst.objectToWorld[4 * row + 3] = static_cast<float>(t[row]);
        }
        inst.materialId = std::string("/inst/mat/i_") + name;
        s.materials.push_back(MixMaterial(inst.materialId, rng));
        s.instances.push_back(std::move(inst));
    }
    AddSkyAndLights(s, "/inst", domeWidth, lights, extent, spacing, rng);
    std::sort(s.meshes.begin(), s.meshes.end(), [](const SceneMesh& a, const SceneMesh& b) { return a.id < b.id; });
    std::sort(s.materials.begin(), s.materials.end(),
              [](const SceneMaterial& a, const SceneMaterial& b) { return a.id < b.id; });
    std::sort(s.lights.begin(), s.lights.end(), [](const SceneLight& a, const SceneLight& b) { return a.id < b.id; });
    std::sort(s.instances.begin(), s.instances.end(),
              [](const SceneInstance& a, const SceneInstance& b) { return a.id < b.id; });
    out.eye = {0.0, extent * 0.45, extent * 0.9};
    out.target = {0.0, spacing, 0.0};
    out.up = {0.0, 1.0, 0.0};
    out.verticalFovDegrees = 45.0;
    return out;
}

} // namespace hdcx
