/*
 * fj_mi355x_ext.h — extensions of the MI355X render library beyond the
 * drop-in boundary of juicer_cuda_api.h: stage-level entry points the
 * reference keeps internal (grain, diffusion, highlight boost, glare), batch
 * rendering and diagnostics, and the host preparation of the print route's
 * tables, which the reference runs in its Rust host (each declaration cites
 * the reference function it computes).
 */
#ifndef FJ_MI355X_EXT_H
#define FJ_MI355X_EXT_H

#include "juicer_cuda_api.h"
#include "juicer_cuda_frame_values.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One timed stage of the last traced frame. A frame is traced when its
 * FjFrame.flags carry FJ_FRAME_TRACE_VERBOSE: every stage is bracketed by HIP
 * events on the frame's stream. */
typedef struct FjxStageTiming {
    char name[32];
    float milliseconds;
} FjxStageTiming;

/* Waits for the last traced frame on `device_id` and copies up to `capacity`
 * stage timings into `out`. Returns the number of stages, or -1 on error. */
int32_t fjx_stage_timings(FjCuda* cuda, int32_t device_id, FjxStageTiming* out, int32_t capacity);

/* Applies the visual-grain stage of `prepared` (its FjVisualGrain, static
 * noise and development sub-layer curves) to three C/M/Y film-density planes
 * in device memory, width x height each (the grain render extent), on the HIP
 * stream `stream` (0: the null stream). This is the stage fj_cuda_render runs
 * between film development and the print exposure
 * (juicer_cuda_apply_visual_grain, src/Cuda/Scan/JuicerCudaScanPipeline.cu:
 * 3141-3756); it is exported for callers that hold densities of their own and
 * for stage-level tests. Returns the status; the planes are modified when the
 * stream completes. */
FjStatus fjx_grain_apply(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* prepared, float* density_c,
                         float* density_m, float* density_y, int32_t width, int32_t height, uint64_t stream,
                         FjErrorBuffer* error);

/* Renders `count` frames of one device in one call (no reference
 * counterpart: the reference renders one frame per fj_cuda_render). Each
 * frame keeps fj_cuda_render's semantics and validation; frame i runs on
 * internal worker stream i % streams (1..4) with its own workspace, so that
 * consecutive frames overlap on the GPU; all are ordered after the work
 * queued on the frames' stream before the call and the stream continues after
 * all of them (every frame must name the same stream, or none: then the call
 * synchronizes). outcomes[i] receives frame i's outcome; the return value is
 * the first failing frame's status, with its index in the error text. */
FjStatus fjx_render_batch(FjCuda* cuda, const FjCudaContext* context, int32_t count, const FjFrame* frames,
                          const FjSubmission* submissions, const FjPreparedHostData* const* prepared, int32_t streams,
                          FjAbortCallback abort_callback, FjRenderOutcome* outcomes, FjErrorBuffer* error);

/* One camera or enlarger diffusion stage (launch_stage,
 * src/Cuda/Diffusion/JuicerCudaDiffusion.cu:779-806, which the reference keeps
 * internal) applied in place to three caller-owned R, G, B float planes of a
 * width x height frame, asynchronously on `stream`: out = (1 - f) x + f (psf
 * conv reflect101(x)). For callers with planes of their own and stage-level
 * tests. */
FjStatus fjx_diffusion_apply(FjCuda* cuda, int32_t device_id, const FjDiffusion* stage, float* red, float* green,
                             float* blue, int32_t width, int32_t height, uint64_t stream, FjErrorBuffer* error);

/* The camera-carrier highlight boost (launch_highlight_boost,
 * src/Cuda/Film/JuicerCudaHighlightBoost.cu:195-205, internal in the
 * reference) applied in place to three caller-owned carrier planes of
 * `samples` floats each in device memory, on `stream`: the carrier maximum and
 * validity, the disposition, then the boost of every sample above the
 * protection. The call waits for the disposition and returns
 * UnsupportedInput "HighlightBoostInvalidCarrier" (a non-finite or negative
 * sample) or "HighlightBoostPeakOverflow" (maximum x gain above FLT_MAX) as
 * check_highlight_boost does (src/ProcessRoot.cpp:4063-4096); the planes
 * are unchanged then. The descriptor needs a recipe hash, finite fields,
 * protection >= 0, sharpness in [0, 28] and gain >= 1. */
FjStatus fjx_highlight_boost_apply(FjCuda* cuda, int32_t device_id, const FjHighlightBoost* boost, float* red,
                                   float* green, float* blue, int64_t samples, uint64_t stream, FjErrorBuffer* error);

/* No counterpart in the reference: fjx_highlight_boost_apply for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, validation,
 * kernels and results as fjx_highlight_boost_apply; its scratch (reduction
 * partials, disposition state, error word: 8,704 bytes) lives in a fixed
 * device array that each GPU holds until the process ends, so the call makes
 * no hipMalloc/hipFree and uses no device state of the owner. Planes of more
 * than 2^34 samples are refused. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race (this call waits for its stream
 * before returning, so calls through one owner never overlap). */
FjStatus fjx_highlight_boost_apply_static(FjCuda* cuda, int32_t device_id, const FjHighlightBoost* boost, float* red,
                                          float* green, float* blue, int64_t samples, uint64_t stream,
                                          FjErrorBuffer* error);

/* The scanner-glare field of one frame window on caller device memory: the
 * reference's optics_glare_generate_kernel and blur_plane_in_place
 * (src/Cuda/Scan/JuicerCudaScanPipeline.cu:273-314, :2361-2405; internal
 * there). Pixel (x, y) of the width x height window draws the Philox4x32-10
 * stream of `seed` (the frame's FNV-1a of recipe hash, frame token and
 * source origin, JuicerCudaExecutor.cpp:2416-2420) at counter
 * {origin_x + x, origin_y + y, medium, n} (medium 1 print, 0 film) and gets
 * the lognormal glare percent of the descriptor's percent and roughness;
 * the field is blurred when FJ_POST_GLARE and a positive blur sigma ask for
 * it. `uniforms` (2 x width x height floats, optional) receives the two raw
 * FP32 uniforms per pixel. Waits for the stream before returning. */
FjStatus fjx_glare_field(FjCuda* cuda, int32_t device_id, const FjScannerPostEffects* post, uint64_t seed,
                         uint32_t medium, int32_t origin_x, int32_t origin_y, int32_t width, int32_t height,
                         float* field, float* uniforms, uint64_t stream, FjErrorBuffer* error);

/* Film TC-LUT build input: the reference's FjFilmTcLutInput
 * (tests/ffi/juicer_test_values.h:964-973; layout identical). Spans are host
 * memory: spectra 192 x 192 x 81 (C, M, wavelength), sensitivity_rgb 81 x 3,
 * reference_illuminant 81, surface_rgb 3 x 15 (read when method is Hanatos
 * and apply_surface is 1), hull_xy 1025 x 2 (read when compression_active and
 * hull_available). method: 0 Hanatos, 2 Arctic. */
typedef struct FjFilmTcLutInput {
    FjFloatSpan spectra, sensitivity_rgb, reference_illuminant;
    float projection_white_xyz[3];
    float spectral_blur;
    uint32_t method, apply_surface;
    FjFloatSpan surface_rgb;
    uint32_t compression_active, hull_available;
    float hull_center_xy[2];
    FjFloatSpan hull_xy;
} FjFilmTcLutInput;

/* Builds the film TC LUT on the GPU (the reference's host build
 * build_tc_lut, rust/film-juicer-core/src/reconstruction.rs:594-677, reached
 * through fj_test_reconstruction_tc_lut): optional spectral blur, the
 * sensitivity integration (Arctic: relit and divided by the projected white's
 * neutral response; Hanatos: optional surface correction) and the
 * input-gamut remap, in the reference's FP32 operations and order, so the
 * 147456 words (R, G, B, 0 per cell) equal the reference's. `samples` is host
 * memory of `count` >= 147456 floats; waits for `stream`. Failures: the
 * reference's preparation failures and diagnostics (TcError), unsupported
 * input for malformed spans or tags (the table is then left untouched;
 * after a preparation failure it is zeroed). */
FjStatus fjx_build_tc_lut(FjCuda* cuda, int32_t device_id, const FjFilmTcLutInput* input, float* samples,
                          size_t count, uint64_t stream, FjErrorBuffer* error);

/* Diagnostic with no counterpart in the reference: arms a one-shot copy of
 * the three C/M/Y film-density planes entering the visual-grain stage of the
 * next frame rendered on `device_id` (the developed density of the grain
 * window, row-major, width x height floats each) into device buffers of at
 * least `samples` floats, on that frame's stream. A frame larger than
 * `samples` disarms the capture without copying; null planes disarm it. Used
 * to check the grain stage on a frame's own densities. */
FjStatus fjx_capture_grain_input(FjCuda* cuda, int32_t device_id, float* density_c, float* density_m,
                                 float* density_y, int64_t samples, FjErrorBuffer* error);

/* The spatial-DIR, visual-grain and film-effects descriptors of one frame,
 * composed from the host's frame math without a device: the composition
 * FjCudaCallAccess.frame_descriptors performs for an inspected frame
 * (native/juicer_cuda_call_scope.cpp:425-551), for callers outside a call
 * scope and for host-only tests. Uses the frame's render window, full-frame
 * extent, pixel size, session seed, clip token and effects geometry only. */
FjStatus fjx_frame_descriptors(uint32_t route, const FjFrame* frame, const FjFrameMathInput* input,
                               FjNativeFrameDescriptors* out, FjErrorBuffer* error);

/* The library's own implementations of the two FjHostPreparation callbacks
 * (juicer_cuda_call_scope.h; the reference's host supplies them from Rust:
 * native/juicer_host_preparation.cpp:28-109 validates what they return), with
 * the callbacks' signatures; `user` is unused. fjx_prepare_scanner_lut builds
 * the PCHIP scanner LUT of rust scanner/lut.rs:186-299 (f32 words equal to the
 * reference's); fjx_prepare_gaussian the normalized FIR weights of rust
 * optics/gaussian.rs:18-56 (radius 1..2048, finite positive sigma). On
 * success the view borrows storage of the calling thread, valid until that
 * thread's next call of the same function; on failure the view is cleared and
 * `failure` carries the reason (FJ_SCANNER_LUT_REASON_*; Gaussian: 1 raw
 * arguments, 2 radius, 3 sigma, 4 allocation, 5 internal). No device is used.
 * A host without Rust preparation can pass these in an FjHostPreparation. */
FjStatus fjx_prepare_scanner_lut(void* user, const FjScannerLutInput* input, FjScannerLutView* view,
                                 FjScannerLutFailure* failure, FjErrorBuffer* error);
FjStatus fjx_prepare_gaussian(void* user, const FjGaussianKernelInput* input, FjGaussianKernelView* view,
                              FjGaussianKernelFailure* failure, FjErrorBuffer* error);

/* ---- Print preparation (host) ----------------------------------------------
 * The three 81-sample preparations behind FjPrint's main illuminant, preflash
 * and normalizer, with the reference's checked test ABI (records of
 * tests/ffi/juicer_test_values.h:870-947, entry points of
 * tests/ffi/juicer_test_api.h:662-671, checks and failure envelopes of
 * rust/film-juicer-plugin/src/print_bridge.rs:187-640). Each zeroes its output
 * and sets the failure to reason NONE (sample/channel 0xFFFFFFFF) first;
 * malformed records or spans give FJ_STATUS_UNSUPPORTED_INPUT with reason RAW,
 * a failure of the computation FJ_STATUS_PREPARATION_FAILURE with the
 * reason and the first failing sample or channel, and the diagnostic
 * "print <kind> reason=R sample|channel=N". No device is used. */
typedef struct FjIlluminant {
    float samples[81];
} FjIlluminant;

typedef struct FjPrintIlluminantParams {
    float edges_nm[4];
    float transitions_nm[4];
    float cmy_cc[3];
} FjPrintIlluminantParams;

typedef struct FjPrintIlluminantFailure {
    uint32_t reason;
    uint32_t sample;
} FjPrintIlluminantFailure;

#define FJ_PRINT_ILLUMINANT_REASON_NONE UINT32_C(0)
#define FJ_PRINT_ILLUMINANT_REASON_RAW UINT32_C(1)
#define FJ_PRINT_ILLUMINANT_REASON_SAMPLE UINT32_C(2)
#define FJ_PRINT_ILLUMINANT_REASON_ENERGY UINT32_C(3)
#define FJ_PRINT_ILLUMINANT_REASON_PANIC UINT32_C(4)

typedef struct FjPrintPreflashInput {
    FjFloatSpan film_base_density;
    FjFloatSpan filtered_illuminant;
    FjFloatSpan print_sensitivity_cmy;
} FjPrintPreflashInput;

typedef struct FjPrintPreflashRaw {
    float cmy[3];
} FjPrintPreflashRaw;

typedef struct FjPrintPreflashFailure {
    uint32_t reason;
    uint32_t channel;
} FjPrintPreflashFailure;

#define FJ_PRINT_PREFLASH_REASON_NONE UINT32_C(0)
#define FJ_PRINT_PREFLASH_REASON_RAW UINT32_C(1)
#define FJ_PRINT_PREFLASH_REASON_CHANNEL UINT32_C(2)
#define FJ_PRINT_PREFLASH_REASON_PANIC UINT32_C(3)

typedef struct FjPrintBalanceInput {
    FjFloatSpan film_base_density;
    FjFloatSpan film_channel_density_cmy;
    FjFloatSpan print_sensitivity_rgb;
    FjFloatSpan filtered_main_illuminant;
    float baseline_density_cmy[3];
    float compensated_density_cmy[3];
    uint64_t reference_hash;
    uint32_t normalization_mode;
} FjPrintBalanceInput;

typedef struct FjPrintBalance {
    float baseline_raw_rgb[3];
    float compensated_raw_rgb[3];
    float factor_midgray;
    float factor_midgray_comp;
    float normalizer;
    uint64_t filtered_main_illuminant_hash;
    uint64_t hash;
} FjPrintBalance;

typedef struct FjPrintBalanceFailure {
    uint32_t reason;
    uint32_t channel;
} FjPrintBalanceFailure;

#define FJ_PRINT_BALANCE_REASON_NONE UINT32_C(0)
#define FJ_PRINT_BALANCE_REASON_RAW UINT32_C(1)
#define FJ_PRINT_BALANCE_REASON_BASELINE_RAW UINT32_C(2)
#define FJ_PRINT_BALANCE_REASON_COMPENSATED_RAW UINT32_C(3)
#define FJ_PRINT_BALANCE_REASON_BALANCE UINT32_C(4)
#define FJ_PRINT_BALANCE_REASON_PANIC UINT32_C(5)

/* The dichroic-filtered enlarger illuminant (rust print.rs:591-639, reached
 * through fj_test_print_illuminant_prepare): per sample the C, M and Y
 * erf-edge filters at 380 + 5 s nm, blended with the CC transmittances
 * 10^(-cc/100), times the 81-sample `source`. FP32 throughout, erf as the
 * reference's pinned Rust libm 0.2.15 erff, powf as the platform's. */
FjStatus fjx_prepare_print_illuminant(const FjPrintIlluminantParams* params, FjFloatSpan source,
                                      FjIlluminant* out_illuminant, FjPrintIlluminantFailure* out_failure,
                                      FjErrorBuffer* error);

/* The preflash raw CMY integral (rust print.rs:538-571, reached through
 * fj_test_print_preflash_prepare): FP64 sum over the 81 samples of
 * 10^-base x filtered light (NaN light counts 0) x finite sensitivity,
 * checked finite and non-negative, then narrowed to f32. */
FjStatus fjx_prepare_print_preflash(const FjPrintPreflashInput* input, FjPrintPreflashRaw* out_raw,
                                    FjPrintPreflashFailure* out_failure, FjErrorBuffer* error);

/* The print balance (rust print.rs:435-522, reached through
 * fj_test_print_balance_prepare): FP64 baseline and compensated raw RGB
 * integrals, the mid-grey factors exp(-mean ln raw), the normalizer of
 * `normalization_mode` (0..3; 4..255 fail as an invalid balance, above 255
 * as a raw input) and the FNV-1a identities of the illuminant and balance. */
FjStatus fjx_prepare_print_balance(const FjPrintBalanceInput* input, FjPrintBalance* out_balance,
                                   FjPrintBalanceFailure* out_failure, FjErrorBuffer* error);

/* A profile's parametric density model (rust film-juicer-core/src/profile.rs:
 * 454-525): per channel C/M/Y and layer, a cumulative Gaussian of the log
 * exposure, amplitude x Phi(+-(logE - center) / sigma); FP64 coefficients,
 * [channel][layer]. */
typedef struct FjDensityCurveModel {
    double centers[3][3];
    double amplitudes[3][3];
    double sigmas[3][3];
} FjDensityCurveModel;

typedef struct FjPrintDensityInput {
    FjDensityCurveModel model; /* the print profile's unadjusted model */
    double gamma;              /* print gamma: finite and > 0; 1 leaves the model's bits untouched */
    uint32_t polarity;         /* FJ_POLARITY_NEGATIVE / _POSITIVE of the profile */
    uint32_t reserved;         /* 0 */
    const double* log_exposure; /* the profile's authored exposure axis */
    size_t count;              /* >= 1 samples */
} FjPrintDensityInput;

typedef struct FjPrintDensityFailure {
    uint32_t reason;
    uint32_t index;   /* failing sample, or 0xFFFFFFFF */
    uint32_t channel; /* failing channel, or 0xFFFFFFFF */
    uint32_t layer;   /* failing layer, or 0xFFFFFFFF */
} FjPrintDensityFailure;

#define FJ_PRINT_DENSITY_REASON_NONE UINT32_C(0)
#define FJ_PRINT_DENSITY_REASON_RAW UINT32_C(1)
#define FJ_PRINT_DENSITY_REASON_INVALID_GAMMA UINT32_C(2)
#define FJ_PRINT_DENSITY_REASON_NONFINITE_LAYER UINT32_C(3)
#define FJ_PRINT_DENSITY_REASON_NONFINITE_TOTAL UINT32_C(4)
#define FJ_PRINT_DENSITY_REASON_ZERO_HASH UINT32_C(5)

/* The print density curves of a print gamma (rust profile/sampled.rs:212-276,
 * PrintDensityCurves): the model adjusted for the gamma in FP64 (centres and
 * sigmas divided by it, sigmas floored at 0.05, NaN kept), then sampled at
 * every authored exposure in FP32 with erfc as the reference's pinned Rust
 * libm 0.2.15 erfcf. Writes `totals` (count x 3, exposure-major C/M/Y),
 * optionally `layers` (count x 9, [sample][layer][channel]) and `print_hash`,
 * FNV-1a of the u64 count, the f32-narrowed axis and the totals' raw words.
 * The axis is not admitted here: it is a completed profile's (finite-or-inf,
 * ascending once narrowed); a NaN exposure fails as a non-finite layer.
 * Outputs are zeroed first and written only on success. Failures: unsupported
 * input for malformed records or short outputs (reason RAW); preparation
 * failure for an invalid gamma, a non-finite layer or total (with its sample,
 * channel and layer) or a zero identity. No device is used. */
FjStatus fjx_prepare_print_density(const FjPrintDensityInput* input, float* totals, float* layers, size_t capacity,
                                   uint64_t* print_hash, FjPrintDensityFailure* out_failure, FjErrorBuffer* error);

/* ---- differentiable print route (no counterpart in the reference) --------
 *
 * The print route as a differentiable function of its input RGB and seven
 * parameters, for fitting a recipe to target frames. The route is the
 * reference's chain (Mallett camera exposure, film curves with negative
 * Langmuir DIR donors, the spatial DIR core + tails, 81-wavelength filtered
 * print exposure with preflash, print curves, scanner LUT, colour, OkLch
 * gamut, linear output) with the visual grain held fixed pathwise: the grain
 * stage's density increment is an input (`grain_delta_cmy`, for instance the
 * difference fjx_grain_apply makes on a frame's densities) added where the
 * grain stage adds it, and a constant of the gradient.
 *
 * The parameters act on the nominal record as the reference's recipe does:
 *   exposure_ev     camera exposure compensation (film_exposure.manual_exposure_ev)
 *   print_exposure  print exposure (print.exposure)
 *   filter_y_cc,    enlarger Y and M filtration added to the record's
 *   filter_m_cc     (print.main_cc_cmy, print.rs:113-176): the record's
 *                   filtered illuminant times F(cc0 + x) / F(cc0) per
 *                   channel, F = 1 - (1 - f)(1 - 10^(-cc/100)) of the
 *                   dichroic filter curves f (print.rs:616-646), cc0 the
 *                   record's M and Y; the print balance normalizer is the
 *                   record's
 *   film_gamma      film development gamma (film_development.gamma_rgb, all
 *                   three channels; the curves are read at gamma * log E)
 *   print_gamma     print development gamma: the print curves read at
 *                   gamma * log E, the density model's centres and widths
 *                   divided by gamma (profile/sampled.rs:212-236)
 *   dir_k           the Langmuir donor K relative to the record's (donor
 *                   K = dir_k * the record's), with the DIR-compensated
 *                   curve axes recomputed as film/dir.rs:234-310 does.
 * Records the fit does not cover are refused as unsupported input: a route
 * other than negative print, TC-LUT reconstruction, input decoding, auto
 * exposure, camera optics, active grain (the increment replaces it), film
 * effects, scanner post effects or glare, an encoded output, and a spatial
 * DIR that is off or not negative-donor. */
typedef struct FjFitParams {
    float exposure_ev;
    float print_exposure;
    float filter_y_cc;
    float filter_m_cc;
    float film_gamma;
    float print_gamma;
    float dir_k;
    float reserved; /* 0 */
} FjFitParams;
#define FJ_FIT_PARAM_COUNT 7

/* Floats of the state a forward saves for its backward: nine planes of
 * width x height (log raw exposure, filtered DIR corrections, densities with
 * the grain increment). */
size_t fjx_fit_saved_floats(int32_t width, int32_t height);

/* Forward of one frame on `device_id`, asynchronously on `stream`. Device
 * pointers: `source` and `destination` are packed RGBA float32 rows (4 x
 * width floats per row); `grain_delta_cmy` is three packed C/M/Y planes of
 * width x height or NULL (no grain); `saved` holds fjx_fit_saved_floats.
 * The destination receives the linear output RGB and the source alpha. */
FjStatus fjx_fit_forward(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* nominal, const FjFitParams* params,
                         const float* source, int32_t width, int32_t height, const float* grain_delta_cmy,
                         float* destination, float* saved, uint64_t stream, FjErrorBuffer* error);

/* Backward of one frame: given the forward's `saved` state and the gradient
 * of a loss with respect to the destination (packed RGBA; the alpha lane is
 * ignored), writes the gradient with respect to the source RGB into
 * `grad_source` (packed RGBA, alpha lane 0) and the gradient with respect to
 * the seven parameters into `grad_params` (8 floats of device memory, in
 * FjFitParams order, the eighth 0). The parameters and record must be the
 * forward's. Asynchronous on `stream`. */
FjStatus fjx_fit_backward(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* nominal, const FjFitParams* params,
                          const float* source, int32_t width, int32_t height, const float* saved,
                          const float* grad_destination, float* grad_source, float* grad_params, uint64_t stream,
                          FjErrorBuffer* error);

#ifdef __cplusplus
}
#endif

#endif
