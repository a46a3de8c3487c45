// fj_glare.h — scanner glare: the per-pixel veiling-glare field of the scan
// (device launches).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fjx {

// One frame window's glare field: pixel (x, y) of the window draws from the
// Philox4x32-10 stream keyed by `seed` with counter {origin_x + x,
// origin_y + y, medium, n} (n = 0, 1: the two uniforms of one Box-Muller
// normal) and stores exp(mu + sigma * normal), the lognormal sample of the
// glare percent (0 where not finite).
struct GlareArgs {
    float* out;          // width x height
    int width, height;
    uint64_t seed;
    uint32_t medium;     // 1 print, 0 film
    int origin_x, origin_y;
    float mu, sigma;     // lognormal parameters of the percent
    float* uniforms;     // optional 2 x width x height: the raw uniforms u1, u2
};

hipError_t launch_glare_generate(const GlareArgs& a, hipStream_t stream);

}  // namespace fjx
