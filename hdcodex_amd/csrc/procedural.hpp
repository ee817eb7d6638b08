// Seeded procedural scene generators (the stand-in for the Hydra producer).
#pragma once

#include "../../include/hdcx_scene.hpp"

#include <array>
#include <cstdint>
#include <string>

namespace hdcx {

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

// "cornell", "heightfield", "triangle", "materials".
ProceduralScene MakeProceduralScene(const std::string& name, std::uint64_t seed,
                                    std::uint32_t param, std::uint32_t domeWidth);

// Every closure lobe, light type and texture path of the integrator in one
// small scene (used for CPU/GPU parity of the full shading code).
ProceduralScene MakeMaterialShowcase(std::uint64_t seed);

} // namespace hdcx
