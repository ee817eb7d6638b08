// hdcx::RenderPass — the reference's Hydra render pass
// (HdCodexRenderPass::_Execute, src/hydra/render_pass.cpp:89-240) without
// OpenUSD: the same reset, interactive and progressive rules, over any
// tracer with hdcx::PathTracer's methods (hdcx_path_tracer.hpp) and a
// published scene, writing Float32Vec4 render buffers
// (HdCodexRenderBuffer, src/hydra/render_buffer.cpp:96-130).
//
// A render delegate that owns an hdcx::PathTracer drives one RenderPass per
// Hydra render pass: Execute() once per _Execute call, with the camera
// MakeCamera (:22-43) builds from the render pass state, the delegate's
// render settings and the AOV bindings. When the tracer offers
// RenderUpscaled (hdcx::PathTracer does), the interactive frame is
// upscaled on the device (hdcx_read_film_upscaled) instead of by the host
// loop UpscaleNearest below; both give the same pixels.
#pragma once

#include "hdcx_scene.hpp"

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

namespace hdcx {

// ParseShadingMode (include/hdcodex/core/shading_mode.h:27-40).
inline std::optional<ShadingMode> ParseShadingMode(std::string_view name) noexcept
{
    if (name == "fused" || name == "fused-specialized") return ShadingMode::Fused;
    if (name == "modular" || name == "instrumented") return ShadingMode::Modular;
    if (name == "raster" || name == "raster-preview") return ShadingMode::RasterPreview;
    return std::nullopt;
}

inline bool operator==(const PathTracerCamera& a, const PathTracerCamera& b)
{
    return a.origin == b.origin && a.lowerLeft == b.lowerLeft && a.horizontal == b.horizontal &&
           a.vertical == b.vertical;
}

// UpscaleNearest (src/hydra/render_pass.cpp:45-68): destination pixel
// (x, y) copies source pixel (min(sw - 1, x * sw / w), min(sh - 1, y * sh / h))
// in unsigned 32-bit arithmetic.
inline std::vector<float> UpscaleNearest(const std::vector<float>& source, unsigned int sourceWidth,
                                         unsigned int sourceHeight, unsigned int width, unsigned int height)
{
    std::vector<float> result(static_cast<std::size_t>(width) * height * 4U);
    for (unsigned int y = 0; y < height; ++y) {
        const unsigned int sy = std::min(sourceHeight - 1U, y * sourceHeight / height);
        for (unsigned int x = 0; x < width; ++x) {
            const unsigned int sx = std::min(sourceWidth - 1U, x * sourceWidth / width);
            const float* from = source.data() + (static_cast<std::size_t>(sy) * sourceWidth + sx) * 4U;
            std::copy(from, from + 4, result.data() + (static_cast<std::size_t>(y) * width + x) * 4U);
        }
    }
    return result;
}

// The render delegate's settings that _Execute reads (:91-121), unclamped.
struct RenderSettings {
    int samplesPerPixel{128};
    int maxBounces{8};
    int samplesPerUpdate{8};
    std::string shadingMode{"fused"};
    unsigned int version{0};  // GetRenderSettingsVersion()
};

// A Float32Vec4 render buffer (the HdCodexRenderBuffer calls _Execute
// makes: Clear, WriteFloat4, SetConverged).
class RenderBuffer {
public:
    RenderBuffer(unsigned int width, unsigned int height)
        : width_(width), height_(height), display_(static_cast<std::size_t>(width) * height * 4U, 0.0F)
    {
    }
    [[nodiscard]] unsigned int GetWidth() const noexcept { return width_; }
    [[nodiscard]] unsigned int GetHeight() const noexcept { return height_; }
    [[nodiscard]] const std::vector<float>& Pixels() const noexcept { return display_; }
    [[nodiscard]] bool IsConverged() const noexcept { return converged_; }
    void SetConverged(bool converged) noexcept { converged_ = converged; }
    void Clear(const std::array<float, 4>& value)
    {
        for (std::size_t i = 0; i < display_.size(); i += 4) std::copy(value.begin(), value.end(), display_.begin() + i);
    }
    // WriteFloat4 (render_buffer.cpp:114-130): sizes must match exactly.
    void WriteFloat4(const std::vector<float>& rgba)
    {
        if (rgba.size() != display_.size()) throw std::invalid_argument("render buffer input does not match dimensions");
        std::memcpy(display_.data(), rgba.data(), rgba.size() * sizeof(float));
    }

private:
    unsigned int width_, height_;
    std::vector<float> display_;
    bool converged_{false};
};

// One AOV binding of the render pass state: the color AOV is rendered, the
// others are cleared when the scene changes.
struct AovBinding {
    RenderBuffer* buffer{nullptr};
    bool color{false};
    std::array<float, 4> clearValue{0.0F, 0.0F, 0.0F, 1.0F};
};

namespace detail {
template <class T, class = void>
struct HasRenderUpscaled : std::false_type {};
template <class T>
struct HasRenderUpscaled<T, std::void_t<decltype(std::declval<T&>().RenderUpscaled(
                                std::declval<const PathTracerCamera&>(), 1U, 1U, 0U, 2U, 1U, 1U, 1U))>>
    : std::true_type {};
} // namespace detail

// Scene: the read side of the reference's VersionedScene
// (include/hdcodex/core/versioned_scene.h): PublishedRevision() and
// Snapshot() returning std::shared_ptr<const SceneSnapshot>.
template <class Tracer, class Scene>
class RenderPass {
public:
    RenderPass(Scene* scene, Tracer* tracer) : scene_(scene), tracer_(tracer) {}

    // IsConverged (:84-87).
    [[nodiscard]] bool IsConverged() const { return converged_ && lastRevision_ == scene_->PublishedRevision(); }
    [[nodiscard]] unsigned int SampleIndex() const noexcept { return sampleIndex_; }
    // The TF_WARN messages of the calls so far.
    [[nodiscard]] const std::vector<std::string>& Warnings() const noexcept { return warnings_; }

    // _Execute (:89-240).
    void Execute(const PathTracerCamera& camera, const RenderSettings& settings, const std::vector<AovBinding>& aovs)
    {
        const unsigned int settingsVersion = settings.version;
        const unsigned int targetSamples = static_cast<unsigned int>(std::clamp(settings.samplesPerPixel, 1, 4096));
        const unsigned int productionBounces = static_cast<unsigned int>(std::clamp(settings.maxBounces, 1, 12));
        const unsigned int samplesPerUpdate = static_cast<unsigned int>(std::clamp(settings.samplesPerUpdate, 1, 64));
        const std::optional<ShadingMode> parsed = ParseShadingMode(settings.shadingMode);
        const ShadingMode shadingMode = parsed.value_or(ShadingMode::Fused);
        if (!parsed) Warn("hdCodex shading mode '" + settings.shadingMode + "' is unknown; using fused");
        const std::uint64_t revision = scene_->PublishedRevision();

        RenderBuffer* color = nullptr;
        const AovBinding* colorBinding = nullptr;
        for (const AovBinding& b : aovs) {
            if (!b.buffer) continue;
            b.buffer->SetConverged(false);
            if (b.color) {
                color = b.buffer;
                colorBinding = &b;
            } else if (lastRevision_ != revision) {
                b.buffer->Clear(b.clearValue);
            }
        }
        if (!color || !tracer_) {
            for (const AovBinding& b : aovs)
                if (b.buffer) b.buffer->SetConverged(true);
            converged_ = true;
            lastRevision_ = revision;
            return;
        }

        const unsigned int width = color->GetWidth();
        const unsigned int height = color->GetHeight();
        if (shadingMode == ShadingMode::RasterPreview) {
            if (lastSettingsVersion_ != settingsVersion)
                Warn("hdCodex raster-preview shading was selected, but its graphics backend is not implemented; "
                     "no compute fallback will be rendered");
            color->Clear(colorBinding->clearValue);
            sampleIndex_ = targetSamples;
            converged_ = true;
            for (const AovBinding& b : aovs)
                if (b.buffer) b.buffer->SetConverged(true);
            Remember(camera, width, height, revision, settingsVersion);
            return;
        }
        const bool cameraChanged = hasCamera_ && !(camera == lastCamera_);
        const bool reset = lastRevision_ != revision || !hasCamera_ || lastSettingsVersion_ != settingsVersion ||
                           cameraChanged || width != lastWidth_ || height != lastHeight_;
        if (reset) {
            try {
                tracer_->SetShadingMode(shadingMode);
                if (lastRevision_ != revision) tracer_->SetScene(scene_->Snapshot());
                sampleIndex_ = 0;
            } catch (const std::exception& error) {
                Warn(std::string("hdCodex failed to build the path-tracing scene: ") + error.what());
                color->Clear(colorBinding->clearValue);
            }
        }

        const bool interactive = cameraChanged && lastRevision_ == revision && width == lastWidth_ &&
                                 height == lastHeight_;
        if (tracer_->HasGeometry() && width > 0 && height > 0 && interactive) {
            try {
                constexpr unsigned int kInteractiveScale = 2U;
                constexpr unsigned int kInteractiveBounces = 2U;
                const unsigned int traceWidth = std::max(1U, (width + kInteractiveScale - 1U) / kInteractiveScale);
                const unsigned int traceHeight = std::max(1U, (height + kInteractiveScale - 1U) / kInteractiveScale);
                if constexpr (detail::HasRenderUpscaled<Tracer>::value) {
                    color->WriteFloat4(tracer_->RenderUpscaled(camera, traceWidth, traceHeight, 0U, kInteractiveBounces,
                                                               1U, width, height));
                } else {
                    color->WriteFloat4(UpscaleNearest(
                        tracer_->Render(camera, traceWidth, traceHeight, 0U, kInteractiveBounces), traceWidth,
                        traceHeight, width, height));
                }
            } catch (const std::exception& error) {
                Warn(std::string("hdCodex interactive path trace failed: ") + error.what());
                color->Clear(colorBinding->clearValue);
            }
        } else if (tracer_->HasGeometry() && width > 0 && height > 0 && sampleIndex_ < targetSamples) {
            try {
                const unsigned int sampleCount = std::min(samplesPerUpdate, targetSamples - sampleIndex_);
                color->WriteFloat4(tracer_->Render(camera, width, height, sampleIndex_, productionBounces, sampleCount));
                sampleIndex_ += sampleCount;
            } catch (const std::exception& error) {
                Warn(std::string("hdCodex path trace failed: ") + error.what());
                color->Clear(colorBinding->clearValue);
                sampleIndex_ = targetSamples;
            }
        } else if (!tracer_->HasGeometry()) {
            color->Clear(colorBinding->clearValue);
            sampleIndex_ = targetSamples;
        }

        converged_ = sampleIndex_ >= targetSamples;
        for (const AovBinding& b : aovs)
            if (b.buffer) b.buffer->SetConverged(converged_);
        Remember(camera, width, height, revision, settingsVersion);
    }

private:
    void Warn(std::string message) { warnings_.push_back(std::move(message)); }
    void Remember(const PathTracerCamera& camera, unsigned int width, unsigned int height, std::uint64_t revision,
                  unsigned int settingsVersion)
    {
        lastCamera_ = camera;
        lastWidth_ = width;
        lastHeight_ = height;
        hasCamera_ = true;
        lastRevision_ = revision;
        lastSettingsVersion_ = settingsVersion;
    }

    Scene* scene_;
    Tracer* tracer_;
    PathTracerCamera lastCamera_{};
    unsigned int lastWidth_{0}, lastHeight_{0};
    bool hasCamera_{false};
    std::uint64_t lastRevision_{~std::uint64_t{0}};
    unsigned int lastSettingsVersion_{~0U};
    unsigned int sampleIndex_{0};
    bool converged_{false};
    std::vector<std::string> warnings_;
};

} // namespace hdcx
