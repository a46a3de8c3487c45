// fj_spatial.h — DCT-domain spatial filters (host orchestration).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_fft.h"
#include "fj_plan.h"

namespace fjx {

// Per-axis eigenvalues of a separable-sum filter: sum_c weights[c] * rows[c][kx] * cols[c][ky].
struct HostTransfer {
    std::vector<float> weights;
    std::vector<std::vector<float>> rows;  // length = width (x frequencies)
    std::vector<std::vector<float>> cols;  // length = height (y frequencies)
};

struct DctFilterJob {
    int width = 0, height = 0, planes = 0;
    const float* in[3] = {};
    float* out[3] = {};
    float* tmp_a[3] = {};
    float* tmp_b[3] = {};
    DctTransfer transfer;
};

std::vector<float> gaussian_weights(int radius, float sigma_px);
void fir_eigenvalues(const std::vector<float>& w, int n, float* out);
void yvv_eigenvalues(double b, const double a[3], int n, float* out);
bool dir_transfer(const FjSpatialDir& sd, int width, int height, HostTransfer& out, std::string& diag);
hipError_t fft_tables(DeviceState& dev, int n, DctRowsArgs& a);
hipError_t upload_transfer(DeviceState& dev, const HostTransfer& h, uint64_t key, DctTransfer& out);
hipError_t run_dct_filter(DeviceState& dev, const DctFilterJob& job, hipStream_t stream);

}  // namespace fjx
