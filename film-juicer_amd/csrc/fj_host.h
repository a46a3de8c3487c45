// fj_host.h — internal host-side declarations of the MI355X render library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fj/juicer_cuda_api.h"
#include "../../include/fj/juicer_cuda_call_scope.h"
#include "fj_device.h"

namespace fjx {

struct ScannerLutInput {
    const float *eps_c, *eps_m, *eps_y;
    const float *weighted_x, *weighted_y, *weighted_z;
    const float* baseline;  // may be null
    int sample_count;
    float inv_yn;
    float range_min_cmy[3];
    float inverse_span_cmy[3];
    int medium;             // 0 film (direct routes), 1 print
    int resolution;
};

struct ScannerLut {
    int resolution = 0;
    std::vector<float> log2_xyz, slope_c, slope_m, slope_y, cell_min, cell_max;
};

// Builds the PCHIP scanner LUT; a non-finite node reports its flat node index
// ((c * r + m) * r + y) and channel.
bool build_scanner_lut(const ScannerLutInput& in, ScannerLut& out, std::string& diagnostic,
                       uint64_t* fail_node = nullptr, uint32_t* fail_channel = nullptr);

// The frame's host preparation (FjHostPreparation of a call scope's render):
// while a scope object lives on a thread, kernel_weights and
// frame_scanner_lut use the host's callbacks when it supplied them, and a
// host failure is kept for the frame to report (take_host_failure).
class HostPreparationScope {
public:
    explicit HostPreparationScope(const FjHostPreparation* host);
    ~HostPreparationScope();
    HostPreparationScope(const HostPreparationScope&) = delete;
    HostPreparationScope& operator=(const HostPreparationScope&) = delete;

private:
    const FjHostPreparation* previous_;
};
bool take_host_failure(uint32_t& category, std::string& text);
// Normalized Gaussian FIR weights (empty on failure).
std::vector<float> kernel_weights(int radius, float sigma_px);
bool frame_scanner_lut(const ScannerLutInput& in, ScannerLut& out, std::string& diagnostic);

// kernels (fj_pixel.hip)
// Persistent launch: each kernel gets the blocks that are resident at once on
// compute_units CUs (at most one per 256-pixel tile).
hipError_t launch_pixel_stage(const PixelParams& p, int mode, int compute_units, hipStream_t stream);
hipError_t launch_post(const PostParams& p, hipStream_t stream);
// Largest unsharp radius the fused post kernel holds in its LDS tile.
int max_post_radius();
hipError_t launch_clear_word(unsigned int* word, hipStream_t stream);
size_t fused_lds_bytes(const PixelParams& p);

}  // namespace fjx
