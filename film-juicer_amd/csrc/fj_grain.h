// fj_grain.h — visual grain in density space (device blocks).
//
// Per channel and size scale, every pixel draws Poisson particle counts from
// its STBN stream (one splitmix64-addressed byte of the blue-noise volume per
// uniform, offset per pixel by a Wang-tile warp), optionally per emulsion
// sub-layer through the layer density curves; the scale's delta against the
// developed density is blurred by a correlation Gaussian and weighted into the
// channel's delta, which is finally added to the density
// (juicer_cuda_apply_visual_grain, JuicerCudaScanPipeline.cu:3141-3756).
//
// The sampler's counts are discontinuous in their inputs, so every operation
// on the decision path is plain FP32 without contraction (the kernels are
// built with contraction off) and exp/log/sin/cos are the deterministic
// Cody–Waite + Taylor forms of fj_detmath.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fjx {

constexpr int kGrainCurveMax = 256;  // sub-layer curves staged in LDS

struct GrainDev {
    // static noise
    const uint8_t* stbn;
    int stbn_w, stbn_h, stbn_frames;
    int stbn_mask;                 // both sides powers of two
    int stbn_off_x, stbn_off_y, stbn_frame;
    long long frame_index;
    const uint8_t* wang_tiles;
    const uint8_t* wang_lut;
    int wang_w, wang_h, wang_count, wang_colors;
    float wang_cell_mm, pixel_size_um;
    uint64_t session_seed, clip_token;
    int origin_x, origin_y, pitch_px;
    float time_alpha;
    // particle model of one pass
    uint64_t seed_base, seed_base_next;
    float density_min[3], density_max[3], n_particles[3], od_particle[3], uniformity[3];
    float dmin_l[3][3], dmax_l[3][3], n_l[3][3], od_l[3][3];
    int n_sublayers, positive_film;
    const float* base_curve[3];    // axis of the layer interpolation (development density, channel c)
    const float* layer_curve[3][3];
    int curve_n;
    int axis_finite[3];
    float axis_prefix[3][16];
    // frame uniforms
    int breathing_active, clump_active;
    float breathing_amp, breathing_t, breathing_cell_small_px, breathing_cell_large_px, breathing_drift_x,
        breathing_drift_y, breathing_roll_y, breathing_mix;
    uint64_t breathing_seed_a, breathing_seed_b, clump_seed;
    float clump_cell_px, clump_roll_y, clump_rms_norm, clump_mu, clump_sigma;
};

enum GrainSampleMode : int { kGrainSimple = 0, kGrainTriplet = 1, kGrainLayer = 2 };

struct GrainSampleArgs {
    GrainDev g;
    const float* density;   // developed density of the channel
    float* out;             // sample plane, or the formed delta when form_delta
    int channel, layer;     // layer for kGrainLayer
    int width, height;
    int form_delta;         // write (sample + bias) - density instead of the sample
    float bias;
};

// Separable half-sample-reflect FIR passes.
struct GrainBlurArgs {
    const float* in = nullptr;
    float* out = nullptr;
    const float* weights = nullptr;  // 2r+1
    int radius = 0;
    int width = 0, height = 0;
    // vertical pass epilogue: out = prev + [weight *] blur, then optionally
    // ((out + bias) - sub) and a final gain
    int weighted = 0, initialize = 0;
    float weight = 0.0f;
    const float* sub = nullptr;
    float bias = 0.0f;
    float gain = 1.0f;               // 1: none
    // vertical pass, unsharp combine instead of the epilogue above:
    // out = x + weight * (x - blur), with x the previous out
    int unsharp = 0;
};

// Elementwise passes over n samples.
enum GrainEltOp : int {
    kEltClear = 0,        // a = 0
    kEltAccumulate,       // a = san((init ? 0 : a) + b)
    kEltAccWeighted,      // a = san((init ? 0 : a) + w b)
    kEltFormDelta,        // a = san((san(a + bias)) - b)
    kEltSubtract,         // a = san((a - b) * 1)
    kEltScale,            // a = san(a * s)
    kEltMixShared,        // a = san(amp (sw shared + iw ind))
    kEltReconstruct,      // a = san((add_mean ? a : 0) + s b)
    kEltCopy,             // a = b
    kEltMixReconstruct,   // kEltMixShared into a temporary m (ind = b), then a = san((add_mean ? a : 0) + s m)
};

struct GrainEltArgs {
    int op = kEltClear;
    float* a = nullptr;
    const float* b = nullptr;
    const float* shared = nullptr;
    long long n = 0;
    int init = 0, add_mean = 0, use_ind = 0;
    float w = 0.0f, bias = 0.0f, s = 0.0f, sw = 0.0f, iw = 0.0f, amp = 0.0f;
};

struct GrainAvgArgs {
    float* d[3];
    long long n;
    float offset, scale;
};

hipError_t launch_grain_samples(const GrainSampleArgs& a, int mode, hipStream_t s);
hipError_t launch_grain_blur_h(const GrainBlurArgs& a, hipStream_t s);
hipError_t launch_grain_blur_v(const GrainBlurArgs& a, hipStream_t s);
hipError_t launch_grain_elt(const GrainEltArgs& a, hipStream_t s);
hipError_t launch_grain_avg3(const GrainAvgArgs& a, hipStream_t s);

}  // namespace fjx
