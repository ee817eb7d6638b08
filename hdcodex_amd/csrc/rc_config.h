// Defaults and checks of a radiance-cache configuration (shared by the
// tracer and the batch renderer).
#pragma once

#include "../../include/hdcx.h"

#include <stdexcept>

namespace hdcx {

inline hdcx_radiance_cache_config NormalizedRcConfig(const hdcx_radiance_cache_config& in)
{
    hdcx_radiance_cache_config c = in;
    if (c.train_samples == 0) c.train_samples = 65536;
    if (c.train_steps == 0) c.train_steps = 4;
    if (!(c.learning_rate > 0.0F)) c.learning_rate = 1e-2F;
    if (!(c.beta1 > 0.0F)) c.beta1 = 0.9F;
    if (!(c.beta2 > 0.0F)) c.beta2 = 0.99F;
    if (!(c.epsilon > 0.0F)) c.epsilon = 1e-8F;
    if (c.terminate_bounce >= HDCX_MAX_BOUNCES) throw std::invalid_argument("terminate_bounce out of range");
    if (!(c.ema_decay >= 0.0F && c.ema_decay < 1.0F)) throw std::invalid_argument("ema_decay outside [0, 1)");
    if (!(c.calibration_decay >= 0.0F && c.calibration_decay < 1.0F))
        throw std::invalid_argument("calibration_decay outside [0, 1)");
    return c;
}

} // namespace hdcx
