// fj_grain_host.h — host side of the visual-grain stage.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_grain.h"
#include "fj_plan.h"

namespace fjx {

struct GrainKernel {
    const float* weights = nullptr;
    int radius = 0;
    float sigma = 0.0f;
};

struct GrainPlan {
    GrainDev base{};
    GrainKernel corr[3];
    GrainKernel dye[3][3];
    float w_fine = 0.0f, w_mid = 0.0f, w_coarse = 0.0f;
    float size_mix_scale = 1.0f, size_mix_gain = 1.0f, amplitude = 0.0f;
    float chroma_mix = 0.0f, shared_w = 0.0f, ind_w = 0.0f, debug_scale = 0.0f;
    int debug_view = 0;
    bool sublayers = false;
};

// decode_grain (native/juicer_cuda_descriptors.cpp:376-397): false names the field.
bool grain_decode(const FjVisualGrain& g, std::string& field);
// Frame binding of an active grain descriptor (native/juicer_cuda_prepared.cpp:200-207).
bool grain_binding(const FjVisualGrain& g, const FjFrame& f);
// Payload and launch preconditions (pack_visual_grain_payload,
// JuicerCudaFilmPayloads.cpp:264-500; juicer_cuda_apply_visual_grain's checks,
// JuicerCudaScanPipeline.cu:3155-3287), device tables and frame uniforms.
// Returns hipSuccess with ok=false and `diag` set when the record is refused.
hipError_t grain_prepare(DeviceState& dev, const FjPreparedHostData& p, GrainPlan& out, bool& ok, std::string& diag,
                         bool& launch_rejected);
// Applies the stage to the three C/M/Y density planes of a width x height
// window; `scratch` holds five planes.
hipError_t run_grain(const GrainPlan& plan, float* const dens[3], int width, int height, float* scratch,
                     hipStream_t stream);
// Frame uniforms (grain_prepare_frame_uniforms_kernel, JuicerCudaFilmPipeline.cu:2174-2352).
void grain_frame_uniforms(const FjVisualGrain& vg, GrainDev& g);

}  // namespace fjx
