// fj_highlight.h — highlight boost of the camera carrier (device launches).
#pragma once

#include <hip/hip_runtime.h>

namespace fjx {

constexpr int kHbMaxBlocks = 1024;
// Largest curve sharpness admitted: the recipe's 28^(1 - range).
constexpr double kHighlightMaxSharpness = 28.0;
// Frame error-word bits of the failing dispositions.
constexpr unsigned int kHbInvalidCarrierBit = 1u << 9;
constexpr unsigned int kHbPeakOverflowBit = 1u << 10;

// Device state of one frame's boost (in the frame's workspace).
struct HbState {
    float partial_max[kHbMaxBlocks];
    unsigned int partial_invalid[kHbMaxBlocks];
    // Per-frame curve constants, prepared in FP64 by one thread and handed
    // to the per-pixel FP32 pass: the protection as a float pair (p = hi +
    // lo), 1 / range, the shape and the boost amplitude delta / g(shape).
    // The exponent's slope shape / range is a float pair too (c = hi + lo).
    float protection_hi, protection_lo, inv_range, slope_hi, slope_lo, amplitude;
    float maximum, peak;
    int applied;
};

struct HbArgs {
    float* planes[3];  // carrier planes, n samples each (boosted in place)
    long long n;
    double protection, sharpness, peak_gain;
    HbState* state;
    unsigned int* error_word;
};

// Reduction, disposition and boost on `stream`.
hipError_t launch_highlight_boost(const HbArgs& a, hipStream_t stream);

}  // namespace fjx
