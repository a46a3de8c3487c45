// fj_autoexposure.h — camera auto-exposure metering (device launches).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_device.h"

namespace fjx {

constexpr int kAeBins = 2048;      // median histogram bins (reference kMedianHistogramBins)
constexpr int kAeMaxBlocks = 256;  // 256-sample blocks of a preview of at most 256 x 256

struct AePartial {
    float sum_y[4];
    float sum_w[4];
    float max_y;
    float pad;
};

// Metering scratch of one frame in flight (device memory).
struct AeScratch {
    AePartial partials[kAeMaxBlocks];
    unsigned int histogram[kAeBins];
    float max_y;
    float scale;  // the metered exposure scale the exposure kernels read
};

struct AeArgs {
    const unsigned char* src;  // the source image at its bounds' origin
    long long row_bytes;
    int comps;
    FjAutoExposure desc;       // preview grid, meter window, method
    ExposureParams ex;         // reconstruction tables; no scale is applied
    AePartial* partials;
    unsigned int* histogram;
    float* max_y;
    float* scale;
};

// Blocks of the partials launch for a descriptor (0: empty preview).
int ae_blocks(const FjAutoExposure& d);
// Meters the preview and writes *a.scale on `stream`.
hipError_t launch_auto_exposure(const AeArgs& a, hipStream_t stream);

}  // namespace fjx
