/*
 * juicer_diffusion_api.h — the diffusion-components import of the render
 * library (replaces native/juicer_diffusion_api.h; records, constants and
 * the function's signature are the reference's, for a host that links it).
 *
 * fj_diffusion_components_prepare is defined by the host (the reference's
 * Rust plug-in: rust/film-juicer-plugin/src/diffusion_bridge.rs:559) and
 * expands a diffusion PSF sample into nine exponential components: lambda in
 * pixels and the red, green and blue weights. The library resolves it at run
 * time when the process defines it and then uses the host's components for
 * each diffusion stage, as the reference's native library does
 * (src/DiffusionHostBehavior.cpp:52-82); without it the library expands the
 * sample itself (rust optics/diffusion_frame.rs:182-220, restated natively).
 *
 * The call is synchronous on the calling thread; inputs stay valid until it
 * returns and no pointer is kept. Each output is cleared before it is
 * written; a failure leaves the components cleared and names reason, stage
 * and field. Diagnostics follow FjErrorBuffer's rules.
 */
#ifndef FJ_JUICER_DIFFUSION_API_H
#define FJ_JUICER_DIFFUSION_API_H

#include "juicer_cuda_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The sample part of an FjDiffusion stage descriptor. */
typedef struct FjDiffusionSample {
    uint32_t family;
    double group_weights[3], group_lambda_um[3];
    double warmth, spatial_scale, pixel_size_um;
    int32_t radius_px;
    uint64_t hash;
} FjDiffusionSample;

/* Nine components: lambda in pixels and per-channel weights. */
typedef struct FjDiffusionComponents {
    double lambda_px[9], red_weights[9], green_weights[9], blue_weights[9];
} FjDiffusionComponents;

typedef struct FjDiffusionFailure {
    uint32_t reason, stage, field;
} FjDiffusionFailure;

#define FJ_DIFFUSION_REASON_NONE 0U
#define FJ_DIFFUSION_REASON_RAW 1U
#define FJ_DIFFUSION_REASON_PREPARATION 2U
#define FJ_DIFFUSION_REASON_PANIC 3U
#define FJ_DIFFUSION_STAGE_NONE 0U
#define FJ_DIFFUSION_STAGE_CAMERA 1U
#define FJ_DIFFUSION_STAGE_ENLARGER 2U
#define FJ_DIFFUSION_FIELD_NONE 0U
#define FJ_DIFFUSION_FIELD_PIXEL_SIZE_UM 1U
#define FJ_DIFFUSION_FIELD_FULL_FRAME_EXTENT 2U
#define FJ_DIFFUSION_FIELD_FAMILY 3U
#define FJ_DIFFUSION_FIELD_RADIUS_PIXELS 4U
#define FJ_DIFFUSION_FIELD_SAMPLE_DESCRIPTOR_HASH 5U
#define FJ_DIFFUSION_FIELD_SAMPLE_DESCRIPTOR_INACTIVE 6U
#define FJ_DIFFUSION_FIELD_GROUP_WEIGHTS 7U
#define FJ_DIFFUSION_FIELD_GROUP_CENTERS_UM 8U
#define FJ_DIFFUSION_FIELD_EFFECTIVE_WARMTH 9U
#define FJ_DIFFUSION_FIELD_SPATIAL_SCALE 10U
#define FJ_DIFFUSION_FIELD_COMPONENT_EXPANSION 11U
#define FJ_DIFFUSION_FIELD_COMPONENT_COUNT 12U
#define FJ_DIFFUSION_FIELD_HALO_CHANNEL_WEIGHTS 13U
#define FJ_DIFFUSION_FIELD_COMPONENT_OUTPUT 14U
#define FJ_DIFFUSION_FIELD_ROUTE 15U
#define FJ_DIFFUSION_FIELD_FULL_FRAME_DOMAIN 16U
#define FJ_DIFFUSION_FIELD_ENLARGER_ROUTE 17U
#define FJ_DIFFUSION_FIELD_LINEAR_DOMAIN 18U
#define FJ_DIFFUSION_FIELD_STAGE_SAMPLE 19U
#define FJ_DIFFUSION_FIELD_STAGE_HASH 20U
#define FJ_DIFFUSION_FIELD_FRAME_SET_HASH 21U

/* Imported (host-defined; resolved at run time, optional). */
FjStatus fj_diffusion_components_prepare(const FjDiffusionSample*, FjDiffusionComponents*, FjDiffusionFailure*,
                                         FjErrorBuffer*);

#ifdef __cplusplus
}
#endif

#endif
