// fj_diffusion_host.h — host side of the camera/enlarger diffusion stages.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_fft.h"
#include "fj_plan.h"

namespace fjx {

constexpr int kDiffusionComponents = 9;

// Nine exponential components of a stage sample: lambda in pixels and the
// R, G, B weights. Returns false and names the rejected field.
bool diffusion_components(const FjDiffusion& d, double lambda_px[kDiffusionComponents],
                          double weights[3][kDiffusionComponents], std::string& field);
// Optics descriptor rules for the diffusion stages; false names the field.
bool diffusion_decode(const FjOptics& o, bool print_route, std::string& field);
// Normalized PSF quadrant q[dy][dx], dx, dy in [0, r], of channel `channel`.
bool diffusion_quadrant(const FjDiffusion& d, int channel, std::vector<double>& q);

struct DiffusionStage {
    int radius = 0;
    int pw = 0, ph = 0;       // padded plane
    float fraction = 0.0f;
    DctTransfer transfer;     // full[c] = T of channel c (R, G, B), pw x ph
};

// Padded size of a stage on a width x height frame: each side extended by the
// radius on both ends, rounded up to the next length with an in-place LDS FFT
// plan (0 when none fits).
void diffusion_padded(const FjDiffusion& d, int width, int height, int& pw, int& ph);
// Scratch floats the transfer build needs.
size_t diffusion_build_floats(const FjDiffusion& d, int pw, int ph);
// Transfer tables of the stage, built on `stream` (cached per PSF and size);
// `scratch` holds diffusion_build_floats() floats that the build may use.
hipError_t diffusion_prepare(DeviceState& dev, const FjDiffusion& d, int width, int height, float* scratch,
                             hipStream_t stream, DiffusionStage& out);
// planes[c] (R, G, B order of the descriptor) are blended in place;
// `padded` holds 6 pw x ph planes. `mark`, when set, is called with 1 just
// before the column filter launch and with 2 just after it (trace events).
hipError_t run_diffusion(DeviceState& dev, const DiffusionStage& st, float* const planes[3], int width, int height,
                         float* padded, hipStream_t stream, const std::function<void(int)>& mark = {});

}  // namespace fjx
