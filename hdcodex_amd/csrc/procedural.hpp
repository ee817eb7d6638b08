// Seeded procedural scene generators (the stand-in for the Hydra producer).
#pragma once

#include "../../include/hdcx_scene.hpp"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace hdcx {

// SplitMix64 (Steele, Lea, Flood 2014).
struct SplitMix64 {
    std::uint64_t state;
    std::uint64_t Next()
    {
        std::uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double Uniform() { return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0); }
    float Range(double lo, double hi) { return static_cast<float>(lo + (hi - lo) * Uniform()); }
};

// Independent stream `stream` of a scene seed.
inline std::uint64_t StreamSeed(std::uint64_t seed, std::uint64_t stream)
{
    SplitMix64 s{seed ^ (stream * 0xd1b54a32d192ed03ULL)};
    return s.Next();
}

// Unit icosphere of `level` subdivisions (20 x 4^level faces), cached.
struct IcoMesh {
    std::vector<std::array<double, 3>> vertices;
    std::vector<std::array<std::uint32_t, 3>> faces;
};
const IcoMesh& UnitIcosphere(int level);
// Appends an icosphere (smooth per-corner normals, spherical UVs) to a mesh.
void AppendIcosphere(SceneMesh& m, int level, const double c[3], double radius);

struct ProceduralScene {
    SceneSnapshot snapshot;
    std::array<double, 3> eye{0.0, 0.0, 0.0};
    std::array<double, 3> target{0.0, 0.0, -1.0};
    std::array<double, 3> up{0.0, 1.0, 0.0};
    double verticalFovDegrees{45.0};
    bool explicitCamera{false};
    PathTracerCamera camera{};

    [[nodiscard]] PathTracerCamera Camera(std::uint32_t width, std::uint32_t height) const;
};

// "cornell", "heightfield", "triangle", "materials", "furnace", "mixed"
// and "mixed_lights" (param: triangle target of the mixed scenes),
// "instanced" (param: instances of a 20,480-triangle sphere, 64 lights) and
// "instanced_small" (param: instances of a 1,280-triangle sphere, 8 lights).
ProceduralScene MakeProceduralScene(const std::string& name, std::uint64_t seed,
                                    std::uint32_t param, std::uint32_t domeWidth);

// Every closure lobe, light type and texture path of the integrator in one
// small scene (used for CPU/GPU parity of the full shading code).
ProceduralScene MakeMaterialShowcase(std::uint64_t seed);

// Icospheres on a ground plane with the six-material mix (Lambert,
// Oren-Nayar, GGX conductor, dispersive dielectric with Abbe U[20,60],
// diffuse translucent, subsurface random walk with mfp U[0.01,0.5]),
// `lights` rect/sphere/disk lights selected by power plus a procedural sky
// dome of domeWidth x domeWidth/2. About triangleTarget triangles.
ProceduralScene MakeMixedScene(std::uint64_t seed, std::uint32_t triangleTarget, std::uint32_t domeWidth,
                               std::uint32_t lights);

// `instances` instances (random TRS, materials from the same mix, one per
// instance) of one level-`level` icosphere prototype (20 x 4^level
// triangles) over a drawn ground quad, with `lights` lights and the dome.
ProceduralScene MakeInstancedScene(std::uint64_t seed, std::uint32_t instances, int level, std::uint32_t domeWidth,
                                   std::uint32_t lights);

} // namespace hdcx
