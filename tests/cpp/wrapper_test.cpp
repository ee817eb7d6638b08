// Exercises include/hdcx_path_tracer.hpp the way the reference's integrator
// test drives VulkanPathTracer (tests/path_tracer_tests.cpp:21-160): build a
// SceneSnapshot in code, SetScene, Render, check the progressive batching
// invariant and basic image properties. Without a device the constructor
// must throw std::runtime_error, which is reported as exit code 3.
#include "hdcx_path_tracer.hpp"

#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>

namespace {
void Check(bool condition, const char* message)
{
    if (!condition) throw std::logic_error(message);
}
}

int main()
{
    std::unique_ptr<hdcx::PathTracer> tracer;
    try {
        tracer = std::make_unique<hdcx::PathTracer>();
    } catch (const std::runtime_error& e) {
        std::printf("no device: %s\n", e.what());
        return 3;
    }
    try {
        Check(tracer->GetShadingMode() == hdcx::ShadingMode::Fused, "default shading mode");
        tracer->SetShadingMode(hdcx::ShadingMode::Modular);
        Check(tracer->GetShadingMode() == hdcx::ShadingMode::Modular, "modular mode");
        bool threw = false;
        try {
            tracer->SetShadingMode(hdcx::ShadingMode::RasterPreview);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        Check(threw, "raster preview must throw invalid_argument");
        tracer->SetShadingMode(hdcx::ShadingMode::Fused);

        auto scene = std::make_shared<hdcx::SceneSnapshot>();
        scene->revision = 1;
        hdcx::SceneMesh mesh;
        mesh.id = "/triangle";
        mesh.positions = {-1.25F, -1.0F, -3.0F, 1.25F, -1.0F, -3.0F, 0.0F, 1.2F, -3.0F};
        mesh.indices = {0, 1, 2};
        mesh.texcoords = {0.0F, 0.0F, 1.0F, 0.0F, 0.5F, 1.0F};
        mesh.materialId = "/red";
        scene->meshes.push_back(mesh);
        hdcx::SceneMaterial red;
        red.id = "/red";
        red.baseColor = {1.0F, 0.1F, 0.1F};
        red.roughness = 0.3F;
        scene->materials.push_back(red);
        hdcx::SceneLight dome;
        dome.id = "/dome";
        dome.color = {0.08F, 0.16F, 0.8F};
        dome.intensity = 0.35F;
        scene->lights.push_back(dome);
        Check(!tracer->HasGeometry(), "no geometry before SetScene");
        Check(tracer->Render({}, 8, 8, 0).empty(), "empty render without geometry");
        tracer->SetScene(scene);
        Check(tracer->HasGeometry(), "geometry after SetScene");
        const hdcx::PathTracerCamera camera{{0.0F, 0.0F, 0.0F}, {-1.0F, -1.0F, -1.0F}, {2.0F, 0.0F, 0.0F},
                                            {0.0F, 2.0F, 0.0F}};
        (void)tracer->Render(camera, 64, 64, 0U);
        const auto two = tracer->Render(camera, 64, 64, 1U);
        const auto batched = tracer->Render(camera, 64, 64, 0U, 8U, 2U);
        Check(two.size() == 64U * 64U * 4U && batched.size() == two.size(), "image size");
        for (std::size_t i = 0; i < two.size(); ++i)
            Check(std::fabs(two[i] - batched[i]) < 1e-5F, "batched samples changed the progressive estimator");
        const auto pixels = tracer->Render(camera, 64, 64, 0U, 8U, 16U);
        for (float v : pixels) Check(std::isfinite(v), "finite output");
        const std::size_t center = (32U * 64U + 32U) * 4U;
        Check(pixels[center] > pixels[center + 1], "red triangle in the centre");
        Check(pixels[center + 3] == 1.0F, "alpha is one");
        Check(tracer->Render(camera, 0, 64, 0U).empty(), "zero width gives an empty image");
        tracer->SetScene(nullptr);
        Check(!tracer->HasGeometry(), "cleared scene");
        std::printf("wrapper ok\n");
        return 0;
    } catch (const std::exception& e) {
        std::printf("FAILED: %s\n", e.what());
        return 1;
    }
}
