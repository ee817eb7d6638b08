// Drives include/hdcx_render_pass.hpp the way Hydra drives the reference's
// HdCodexRenderPass::_Execute (src/hydra/render_pass.cpp:89-240).
//
// Built without HDCX_RENDER_PASS_GPU: a recording tracer stands in for the
// device, and each case checks the calls the pass makes (SetShadingMode,
// SetScene, Render with its sizes, sample index, bounces and sample count),
// what lands in the render buffers, the warnings and the convergence flags,
// following the reference's rules case by case: reset on a new scene
// revision, camera, size or settings version; the interactive
// half-resolution frame (2 bounces, nearest upscale) right after a camera
// move; progressive batches of min(samplesPerUpdate, target - index);
// setting clamps; raster preview; no geometry; failures.
//
// Built with HDCX_RENDER_PASS_GPU (linked with libhdcx): the same pass over
// hdcx::PathTracer on a small scene; every frame it writes equals, bit for
// bit, what a second tracer returns for the same Render calls (and the
// interactive frame the host UpscaleNearest of the half-resolution render,
// while the pass upscales on the device). Exit code 0 = pass, 1 = failure,
// 3 = no device (GPU build only).
#include "hdcx_render_pass.hpp"
#ifdef HDCX_RENDER_PASS_GPU
#include "hdcx_path_tracer.hpp"
#endif

#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

int gFailures = 0;

void Expect(bool ok, const std::string& what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++gFailures;
    }
}

hdcx::PathTracerCamera Camera(float z)
{
    hdcx::PathTracerCamera c;
    c.origin = {0.0F, 0.0F, z};
    c.lowerLeft = {-1.0F, -1.0F, z - 1.0F};
    c.horizontal = {2.0F, 0.0F, 0.0F};
    c.vertical = {0.0F, 2.0F, 0.0F};
    return c;
}

// The read side of a versioned scene: a revision and its snapshot.
struct Scene {
    std::uint64_t revision{1};
    std::shared_ptr<hdcx::SceneSnapshot> snapshot;
    [[nodiscard]] std::uint64_t PublishedRevision() const { return revision; }
    [[nodiscard]] std::shared_ptr<const hdcx::SceneSnapshot> Snapshot() const
    {
        snapshot->revision = revision;
        return snapshot;
    }
};

std::shared_ptr<hdcx::SceneSnapshot> Triangle()
{
    auto scene = std::make_shared<hdcx::SceneSnapshot>();
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
    return scene;
}

const std::vector<float>& Pixels(const hdcx::RenderBuffer& b) { return b.Pixels(); }

bool AllEqual(const std::vector<float>& v, const std::array<float, 4>& value)
{
    for (std::size_t i = 0; i < v.size(); ++i)
        if (v[i] != value[i % 4]) return false;
    return true;
}

#ifndef HDCX_RENDER_PASS_GPU

// Records the calls; a render returns pixel p, channel c = 1e6 * call +
// 1000 * sampleIndex + 4 * p + c, so every frame is distinguishable.
struct RecordingTracer {
    std::vector<std::string> calls;
    bool geometry{false};
    bool failRender{false};
    bool failScene{false};
    int renders{0};

    void SetShadingMode(hdcx::ShadingMode mode)
    {
        calls.push_back(mode == hdcx::ShadingMode::Modular ? "mode modular" : "mode fused");
        if (mode == hdcx::ShadingMode::RasterPreview) throw std::invalid_argument("raster preview");
    }
    void SetScene(const std::shared_ptr<const hdcx::SceneSnapshot>& scene)
    {
        calls.push_back("scene " + std::to_string(scene ? scene->revision : 0));
        if (failScene) throw std::runtime_error("scene build failed");
        geometry = scene && !scene->meshes.empty();
    }
    [[nodiscard]] bool HasGeometry() const { return geometry; }
    std::vector<float> Render(const hdcx::PathTracerCamera&, std::uint32_t w, std::uint32_t h, std::uint32_t index,
                              std::uint32_t bounces = 8U, std::uint32_t count = 1U)
    {
        calls.push_back("render " + std::to_string(w) + "x" + std::to_string(h) + " index " + std::to_string(index) +
                        " bounces " + std::to_string(bounces) + " count " + std::to_string(count));
        if (failRender) throw std::runtime_error("device lost");
        ++renders;
        std::vector<float> out(static_cast<std::size_t>(w) * h * 4U);
        for (std::size_t i = 0; i < out.size(); ++i) out[i] = 1e6F * static_cast<float>(renders) + 1000.0F * index + i;
        return out;
    }
    std::vector<float> Last(std::uint32_t w, std::uint32_t h, std::uint32_t index) const
    {
        std::vector<float> out(static_cast<std::size_t>(w) * h * 4U);
        for (std::size_t i = 0; i < out.size(); ++i) out[i] = 1e6F * static_cast<float>(renders) + 1000.0F * index + i;
        return out;
    }
    std::vector<std::string> Take()
    {
        std::vector<std::string> out;
        out.swap(calls);
        return out;
    }
};

using Pass = hdcx::RenderPass<RecordingTracer, Scene>;

std::string Join(const std::vector<std::string>& v)
{
    std::string s;
    for (const auto& x : v) s += "[" + x + "]";
    return s;
}

void ExpectCalls(RecordingTracer& t, const std::vector<std::string>& want, const std::string& what)
{
    const auto got = t.Take();
    Expect(got == want, what + ": calls " + Join(got) + " expected " + Join(want));
}

void UpscaleKnownAnswers()
{
    // 2x2 -> 3x3: columns and rows 0, 0, 1 (x * 2 / 3).
    std::vector<float> src(16);
    for (int i = 0; i < 16; ++i) src[i] = static_cast<float>(i);
    const auto up = hdcx::UpscaleNearest(src, 2, 2, 3, 3);
    const int map[3] = {0, 0, 1};
    bool ok = true;
    for (int y = 0; y < 3; ++y)
        for (int x = 0; x < 3; ++x)
            for (int c = 0; c < 4; ++c) ok = ok && up[(y * 3 + x) * 4 + c] == src[(map[y] * 2 + map[x]) * 4 + c];
    Expect(ok, "UpscaleNearest 2x2 -> 3x3");
    // Same size: identity. Half size: every other pixel (x * 4 / 2 = 2x).
    Expect(hdcx::UpscaleNearest(src, 2, 2, 2, 2) == src, "UpscaleNearest identity");
    std::vector<float> big(4 * 4 * 4);
    for (std::size_t i = 0; i < big.size(); ++i) big[i] = static_cast<float>(i);
    const auto down = hdcx::UpscaleNearest(big, 4, 4, 2, 2);
    Expect(down[0] == big[0] && down[4] == big[8] && down[8] == big[32] && down[12] == big[40], "UpscaleNearest 4x4 -> 2x2");
}

void Cases()
{
    Scene scene;
    scene.snapshot = Triangle();
    RecordingTracer tracer;
    Pass pass(&scene, &tracer);
    hdcx::RenderBuffer color(8, 6), depth(8, 6);
    const std::array<float, 4> clearColor{0.1F, 0.2F, 0.3F, 1.0F}, clearDepth{1.0F, 1.0F, 1.0F, 1.0F};
    std::vector<hdcx::AovBinding> aovs{{&color, true, clearColor}, {&depth, false, clearDepth}};
    hdcx::RenderSettings settings;
    const auto camA = Camera(0.0F), camB = Camera(0.5F);

    // First frame: reset (shading mode, scene), progressive batch of 8.
    pass.Execute(camA, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "scene 1", "render 8x6 index 0 bounces 8 count 8"}, "first frame");
    Expect(Pixels(color) == tracer.Last(8, 6, 0), "first frame written to the color buffer");
    Expect(AllEqual(Pixels(depth), clearDepth), "other AOVs cleared on a new revision");
    Expect(pass.SampleIndex() == 8 && !pass.IsConverged() && !color.IsConverged(), "first frame not converged");

    // Same camera: batches of 8 up to the 128-sample target, then nothing.
    for (unsigned k = 1; k < 16; ++k) {
        pass.Execute(camA, settings, aovs);
        ExpectCalls(tracer, {"render 8x6 index " + std::to_string(8 * k) + " bounces 8 count 8"}, "progressive batch");
    }
    Expect(pass.SampleIndex() == 128 && pass.IsConverged() && color.IsConverged() && depth.IsConverged(), "converged");
    pass.Execute(camA, settings, aovs);
    ExpectCalls(tracer, {}, "nothing after convergence");

    // Camera move: reset (no scene rebuild), interactive half-resolution
    // frame with 2 bounces, nearest-upscaled; the next call restarts at 0.
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "render 4x3 index 0 bounces 2 count 1"}, "interactive frame");
    Expect(Pixels(color) == hdcx::UpscaleNearest(tracer.Last(4, 3, 0), 4, 3, 8, 6), "interactive frame upscaled");
    Expect(pass.SampleIndex() == 0 && !pass.IsConverged(), "interactive frame resets the samples");
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"render 8x6 index 0 bounces 8 count 8"}, "progressive after the interactive frame");

    // New revision and a camera move together: rebuild, not interactive.
    scene.revision = 2;
    pass.Execute(camA, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "scene 2", "render 8x6 index 0 bounces 8 count 8"}, "new revision");

    // Resize (odd size): reset, not interactive; then a camera move at the
    // new size traces ceil(w / 2) x ceil(h / 2).
    hdcx::RenderBuffer odd(7, 5);
    std::vector<hdcx::AovBinding> oddAovs{{&odd, true, clearColor}};
    pass.Execute(camA, settings, oddAovs);
    ExpectCalls(tracer, {"mode fused", "render 7x5 index 0 bounces 8 count 8"}, "resize");
    pass.Execute(camB, settings, oddAovs);
    ExpectCalls(tracer, {"mode fused", "render 4x3 index 0 bounces 2 count 1"}, "interactive at an odd size");
    Expect(Pixels(odd) == hdcx::UpscaleNearest(tracer.Last(4, 3, 0), 4, 3, 7, 5), "odd-size upscale");

    // Settings: a new version resets; values are clamped (spp 1..4096,
    // bounces 1..12, samples per update 1..64).
    settings.version = 1;
    settings.samplesPerPixel = 100000;
    settings.maxBounces = 99;
    settings.samplesPerUpdate = 1000;
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "render 8x6 index 0 bounces 12 count 64"}, "clamped settings");
    settings.version = 2;
    settings.samplesPerPixel = 20;
    settings.maxBounces = 0;
    settings.samplesPerUpdate = 0;
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "render 8x6 index 0 bounces 1 count 1"}, "lower clamps");
    settings.samplesPerUpdate = 64;
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"render 8x6 index 1 bounces 1 count 19"}, "last batch up to the target");
    Expect(pass.IsConverged(), "converged at 20");

    // Shading modes: modular; an unknown name warns and uses fused.
    settings.version = 3;
    settings.shadingMode = "instrumented";
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode modular", "render 8x6 index 0 bounces 1 count 20"}, "modular");
    const std::size_t warnings = pass.Warnings().size();
    settings.version = 4;
    settings.shadingMode = "bogus";
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "render 8x6 index 0 bounces 1 count 20"}, "unknown shading mode");
    Expect(pass.Warnings().size() == warnings + 1, "unknown shading mode warns");

    // Raster preview: no tracer call, color cleared, converged; the warning
    // once per settings version.
    settings.version = 5;
    settings.shadingMode = "raster-preview";
    pass.Execute(camB, settings, aovs);
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {}, "raster preview");
    Expect(AllEqual(Pixels(color), clearColor) && pass.IsConverged() && color.IsConverged(), "raster preview clears");
    Expect(pass.Warnings().size() == warnings + 2, "raster preview warns once per settings version");

    // A render failure: warning, color cleared, converged (index = target).
    settings.version = 6;
    settings.shadingMode = "fused";
    tracer.failRender = true;
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "render 8x6 index 0 bounces 1 count 20"}, "failed render");
    Expect(AllEqual(Pixels(color), clearColor) && pass.IsConverged(), "failed render clears and converges");
    tracer.failRender = false;

    // A scene-build failure: warning and clear; the index is not reset, so
    // nothing renders while the old scene's index is at the target.
    scene.revision = 3;
    tracer.failScene = true;
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "scene 3"}, "failed scene build");
    Expect(AllEqual(Pixels(color), clearColor), "failed scene build clears");
    tracer.failScene = false;

    // No geometry: cleared, converged.
    scene.revision = 4;
    scene.snapshot = std::make_shared<hdcx::SceneSnapshot>();
    pass.Execute(camB, settings, aovs);
    ExpectCalls(tracer, {"mode fused", "scene 4"}, "empty scene");
    Expect(AllEqual(Pixels(color), clearColor) && pass.IsConverged(), "empty scene clears and converges");

    // No color AOV: converged without tracer calls; IsConverged follows the
    // published revision.
    Pass bare(&scene, &tracer);
    std::vector<hdcx::AovBinding> depthOnly{{&depth, false, clearDepth}};
    bare.Execute(camA, settings, depthOnly);
    ExpectCalls(tracer, {}, "no color AOV");
    Expect(bare.IsConverged() && depth.IsConverged(), "no color AOV converges");
    scene.revision = 5;
    Expect(!bare.IsConverged(), "a newer published revision is not converged");
}

#else

bool SameBits(const std::vector<float>& a, const std::vector<float>& b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

void GpuCases()
{
    Scene scene;
    scene.snapshot = Triangle();
    hdcx::PathTracer tracer, check;
    check.SetScene(scene.Snapshot());
    hdcx::RenderPass<hdcx::PathTracer, Scene> pass(&scene, &tracer);
    const unsigned W = 37, H = 23;
    hdcx::RenderBuffer color(W, H);
    std::vector<hdcx::AovBinding> aovs{{&color, true, {0.0F, 0.0F, 0.0F, 1.0F}}};
    hdcx::RenderSettings settings;
    settings.samplesPerPixel = 16;
    const auto camA = Camera(0.0F), camB = Camera(0.25F);

    pass.Execute(camA, settings, aovs);
    Expect(SameBits(Pixels(color), check.Render(camA, W, H, 0, 8, 8)), "first frame equals Render(0, 8 bounces, 8)");
    pass.Execute(camA, settings, aovs);
    Expect(SameBits(Pixels(color), check.Render(camA, W, H, 8, 8, 8)), "second batch equals Render(8, 8)");
    Expect(pass.IsConverged() && color.IsConverged(), "converged at 16 samples");

    pass.Execute(camB, settings, aovs);
    const auto half = check.Render(camB, 19, 12, 0, 2);
    Expect(SameBits(Pixels(color), hdcx::UpscaleNearest(half, 19, 12, W, H)),
           "interactive frame (device upscale) equals host UpscaleNearest of the half-resolution render");
    Expect(!pass.IsConverged() && pass.SampleIndex() == 0, "interactive frame resets");
    pass.Execute(camB, settings, aovs);
    Expect(SameBits(Pixels(color), check.Render(camB, W, H, 0, 8, 8)), "progressive after the interactive frame");

    // Device upscale against the host loop for other ratios, down-scaling too.
    const unsigned sizes[][4] = {{5, 3, 64, 48}, {64, 48, 5, 3}, {33, 17, 33, 17}, {1, 1, 9, 4}};
    for (const auto& s : sizes) {
        const auto direct = check.Render(camA, s[0], s[1], 0, 2);
        const auto up = check.RenderUpscaled(camA, s[0], s[1], 0, 2, 1, s[2], s[3]);
        Expect(SameBits(up, hdcx::UpscaleNearest(direct, s[0], s[1], s[2], s[3])),
               "device upscale " + std::to_string(s[0]) + "x" + std::to_string(s[1]) + " -> " + std::to_string(s[2]) +
                   "x" + std::to_string(s[3]));
    }
}

#endif

} // namespace

int main()
{
#ifndef HDCX_RENDER_PASS_GPU
    UpscaleKnownAnswers();
    Cases();
#else
    try {
        GpuCases();
    } catch (const std::runtime_error& e) {
        std::printf("no device or device error: %s\n", e.what());
        return 3;
    }
#endif
    if (gFailures) {
        std::printf("%d failures\n", gFailures);
        return 1;
    }
    std::printf("render pass ok\n");
    return 0;
}
