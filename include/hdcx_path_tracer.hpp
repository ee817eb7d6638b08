// hdcx::PathTracer — the reference integrator's C++ method surface
// (class hdcodex::VulkanPathTracer, include/hdcodex/gpu/vulkan_path_tracer.h:26-52)
// on top of the hdcx C ABI. A render pass written against the reference
// (src/hydra/render_pass.cpp:89-240) needs only the type swap: SetScene,
// SetShadingMode, GetShadingMode, Render and HasGeometry keep their
// signatures, defaults, return values and exception types.
#pragma once

#include "hdcx.h"
#include "hdcx_scene.hpp"

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hdcx {

struct PackedScene;

// Scene packing exported by libhdcx (C++ linkage): the reference's
// BuildScene flattening (src/gpu/vulkan_path_tracer.cpp:2604-2993).
class PackedSceneHandle {
public:
    explicit PackedSceneHandle(const SceneSnapshot& snapshot);
    ~PackedSceneHandle();
    PackedSceneHandle(const PackedSceneHandle&) = delete;
    PackedSceneHandle& operator=(const PackedSceneHandle&) = delete;
    [[nodiscard]] const hdcx_scene_view& View() const noexcept { return view_; }

private:
    PackedScene* packed_;
    hdcx_scene_view view_{};
};

// Device selection (the reference passes a VulkanContext; here a HIP device
// ordinal plays that role).
struct DeviceContext {
    int deviceOrdinal{0};
};

class PathTracer final {
public:
    explicit PathTracer(const DeviceContext& context = {})
    {
        if (hdcx_create(context.deviceOrdinal, &tracer_) != HDCX_OK) throw std::runtime_error(hdcx_last_error());
    }
    ~PathTracer() { hdcx_destroy(tracer_); }
    PathTracer(const PathTracer&) = delete;
    PathTracer& operator=(const PathTracer&) = delete;

    // Full rebuild, blocking; throws std::runtime_error (:2604-3128).
    void SetScene(const std::shared_ptr<const SceneSnapshot>& scene)
    {
        if (!scene) {
            Check(hdcx_set_scene(tracer_, nullptr));
            return;
        }
        PackedSceneHandle packed(*scene);
        Check(hdcx_set_scene(tracer_, &packed.View()));
    }

    // RasterPreview throws std::invalid_argument (:2495-2498).
    void SetShadingMode(ShadingMode mode)
    {
        const int code = hdcx_set_shading_mode(tracer_, static_cast<int>(mode));
        if (code == HDCX_ERR_INVALID_ARGUMENT) throw std::invalid_argument(hdcx_last_error());
        Check(code);
    }

    [[nodiscard]] ShadingMode GetShadingMode() const noexcept
    {
        int mode = HDCX_SHADING_FUSED;
        hdcx_get_shading_mode(tracer_, &mode);
        return static_cast<ShadingMode>(mode);
    }

    // Linear RGBA32F, row 0 at the bottom, alpha 1; the running mean of
    // samples [0, sampleIndex + sampleCount); empty without geometry or
    // for a zero size (:3235).
    [[nodiscard]] std::vector<float> Render(const PathTracerCamera& camera, std::uint32_t width, std::uint32_t height,
                                            std::uint32_t sampleIndex, std::uint32_t maxBounces = 8U,
                                            std::uint32_t sampleCount = 1U)
    {
        if (!HasGeometry() || width == 0 || height == 0) return {};
        hdcx_camera c{};
        for (int k = 0; k < 3; ++k) {
            c.origin[k] = camera.origin[k];
            c.lower_left[k] = camera.lowerLeft[k];
            c.horizontal[k] = camera.horizontal[k];
            c.vertical[k] = camera.vertical[k];
        }
        std::vector<float> pixels(static_cast<std::size_t>(width) * height * 4U);
        const int code = hdcx_render(tracer_, &c, width, height, sampleIndex, maxBounces, sampleCount, pixels.data());
        if (code == HDCX_ERR_NO_GEOMETRY) return {};
        Check(code);
        return pixels;
    }

    // Render at traceWidth x traceHeight, then the film upscaled by nearest
    // neighbour to width x height on the device: what the render pass's
    // interactive frame (src/hydra/render_pass.cpp:191-208) computes with
    // Render + UpscaleNearest. Empty as Render is.
    [[nodiscard]] std::vector<float> RenderUpscaled(const PathTracerCamera& camera, std::uint32_t traceWidth,
                                                    std::uint32_t traceHeight, std::uint32_t sampleIndex,
                                                    std::uint32_t maxBounces, std::uint32_t sampleCount,
                                                    std::uint32_t width, std::uint32_t height)
    {
        if (!HasGeometry() || traceWidth == 0 || traceHeight == 0 || width == 0 || height == 0) return {};
        hdcx_camera c{};
        for (int k = 0; k < 3; ++k) {
            c.origin[k] = camera.origin[k];
            c.lower_left[k] = camera.lowerLeft[k];
            c.horizontal[k] = camera.horizontal[k];
            c.vertical[k] = camera.vertical[k];
        }
        const int code = hdcx_render(tracer_, &c, traceWidth, traceHeight, sampleIndex, maxBounces, sampleCount, nullptr);
        if (code == HDCX_ERR_NO_GEOMETRY) return {};
        Check(code);
        std::vector<float> pixels(static_cast<std::size_t>(width) * height * 4U);
        Check(hdcx_read_film_upscaled(tracer_, width, height, pixels.data(), 0));
        return pixels;
    }

    [[nodiscard]] bool HasGeometry() const noexcept { return hdcx_has_geometry(tracer_) != 0; }

    [[nodiscard]] hdcx_tracer* Native() const noexcept { return tracer_; }

private:
    static void Check(int code)
    {
        if (code != HDCX_OK) throw std::runtime_error(hdcx_last_error());
    }
    hdcx_tracer* tracer_{nullptr};
};

} // namespace hdcx
