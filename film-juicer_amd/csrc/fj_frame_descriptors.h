// fj_frame_descriptors.h — native composition of the per-frame descriptors
// (spatial DIR, visual grain, film effects) from the host's frame math, and
// the auto-exposure preview descriptor of the call scope.
#pragma once

#include <string>

#include "../../include/fj/juicer_cuda_frame_values.h"

namespace fjx {

// Spatial-DIR descriptor of the host's component plan
// (compose_spatial_dir_descriptor, src/RenderValues.cpp:279-289, with
// bind_spatial_dir_math :187-232, empty_dir_frame_math :244-274 and
// finish_spatial_dir_descriptor :169-185).
// Absent math must be all zero; then `out` is left empty. False: mismatch.
bool compose_spatial_dir(const FjDirFrameMath& math, float pixel_size_um, FjSpatialDir& out);

// Visual-grain descriptor (compose_visual_grain_frame_descriptor,
// src/FilmEffectsFrameDescriptors.cpp:294-326, bind_visual_grain_math
// :174-279, empty_grain_math :118-172, hash_descriptor :40-112), encoded as
// the render record
// (encode_grain, native/juicer_cuda_call_scope.cpp:109-171).
// capture_polarity: 0 negative, 1 positive, 2 unsupported route.
// False sets `diagnostic`.
bool compose_visual_grain(const FjGrainCoefficients& coefficients, uint32_t capture_polarity,
                          const FjGrainFrameMath& math, FjVisualGrain& out, std::string& diagnostic);

// Absent film-effects math must be all zero
// (compose_frame_descriptor, src/FilmEffectsFrameDescriptors.cpp:451-456).
bool effects_math_empty(const FjEffectsFrameMath& math);

// The three descriptors of one frame (CallScope::frame_descriptors,
// native/juicer_cuda_call_scope.cpp:436-544): each present math must carry the
// frame's render and full-frame extents (grain: and its pixel size, bit for
// bit); DIR and grain compose as above; an active grain descriptor must carry
// the frame's session seed and clip token. Film-effects math must be absent
// (effects are not rendered by this library). Returns the status category
// (FJ_STATUS_SUCCESS, FJ_STATUS_PREPARATION_FAILURE for a mismatch,
// FJ_STATUS_UNSUPPORTED_INPUT for effects); `diagnostic` names the failure.
uint32_t compose_frame_descriptors(uint32_t route, const FjFrame& frame, const FjFrameMathInput& in,
                                   FjNativeFrameDescriptors& out, std::string& diagnostic);

// Auto-exposure preview descriptor of a meter window
// (make_auto_exposure_preview_descriptor, src/Cuda/JuicerCudaExecutor.cpp:774-827):
// the meter extent scaled to a long edge of at most 256, and its identity hash.
void auto_exposure_preview(const FjRect& source, const FjRect& meter, uint32_t method, FjAutoExposure& out);

}  // namespace fjx
