/*
 * fj_oracle.h — FP64 CPU oracle of Film-Juicer's per-frame render path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load it. The product never links or calls it.
 *
 * Every function restates the reference algorithm in double precision from
 * the same prepared FP32 tables the product consumes (FjPreparedHostData);
 * each definition cites the reference file:line it follows.
 */
#ifndef FJ_ORACLE_H
#define FJ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/fj/juicer_cuda_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PCHIP slopes of one line (rust/film-juicer-core/src/scanner/lut.rs:108-138). */
void fjo_pchip_slopes(const double* values, int n, double* out);

/* Scanner PCHIP LUT build (scanner/lut.rs:140-299). Outputs are f32 words:
 * log2_xyz/slope_* hold res^3*3, cell_min/max hold (res-1)^3*3 floats.
 * Returns 0 on success, 1 on a non-finite node (fail_index/fail_channel set),
 * 2 on bad arguments. */
int fjo_scanner_lut_build(const float* eps_c, const float* eps_m, const float* eps_y,
                          const float* weighted_x, const float* weighted_y, const float* weighted_z,
                          const float* baseline_or_null, int sample_count, float inv_yn,
                          const float range_min_cmy[3], const float inverse_span_cmy[3],
                          int medium, int resolution,
                          float* log2_xyz, float* slope_c, float* slope_m, float* slope_y,
                          float* cell_min, float* cell_max,
                          uint64_t* fail_index, uint32_t* fail_channel);

/* Gaussian FIR weights of 2*radius+1 taps (rust optics/gaussian.rs:20-56). */
int fjo_gaussian_weights(int radius, float sigma, float* out);

/* Visual grain (fjo_grain.c; FP32 restatement of
 * juicer_cuda_apply_visual_grain, JuicerCudaScanPipeline.cu:3141-3756): applies
 * the record's grain to three C/M/Y density planes in place. 0 on success. */
int fjo_grain_apply(const FjPreparedHostData* p, float* const dens[3], int width, int height);
/* One uniform of the STBN stream of pixel (x, y) of the render window. */
float fjo_grain_uniform(const FjPreparedHostData* p, uint64_t seed, int x, int y, int frame_offset, uint32_t draw);
uint64_t fjo_splitmix64(uint64_t v);
/* Deterministic FP32 elementary functions of the grain restatement. */
float fjo_det_expf(float x);
float fjo_det_logf(float x);
float fjo_det_sinf(float x);
float fjo_det_cosf(float x);

/* Reflected-YVV boundary solve for a line of n samples: weights[3*n]
 * (start-state rows) and terminal[size*size]; returns size = min(n, 3). */
int fjo_yvv_boundary(const double c[9], int n, double* weights, double* terminal);

/* Reflected third-order recursive filter on one strided line of n doubles
 * (forward + backward pass with the periodic boundary solve of rust
 * film/dir/boundary.rs). c = {b, a1, a2, a3, D, N, n1, n2, n3}. */
void fjo_yvv_reflect_line(const double* in, double* out, int n, ptrdiff_t stride, const double c[9]);

/* Spatial DIR operator on `nplanes` W x H double planes in place. */
int fjo_dir_filter(const FjSpatialDir* sd, double* const* planes, int nplanes, int W, int H);

/* Scatter/halation on three W x H double planes (R, G, B) in place
 * (JuicerCudaScatterHalation.cu:690-852). */
int fjo_halation(const FjScatterHalation* d, double* const planes[3], int W, int H);

/* Diffusion PSF components (rust optics/diffusion_frame.rs:286-420):
 * lambda_px[9] and per-channel (R, G, B) weights. Returns 0 on success. */
int fjo_diffusion_components(const FjDiffusion* d, double lambda_px[9], double weights[3][9]);
/* Normalized f32 PSF samples of one channel, (2r+1)^2 row-major. */
int fjo_diffusion_psf(const FjDiffusion* d, int channel, float* psf);
/* Diffusion stage (whole-sample mirrored borders, (1-f) x + f psf*x) on
 * three W x H planes (R, G, B) in place (JuicerCudaDiffusion.cu:403-487). */
int fjo_diffusion(const FjDiffusion* d, double* const planes[3], int W, int H);

typedef struct FjoPlan FjoPlan;

/* A diffusion stage implementation with fjo_diffusion's contract (the same
 * operator on the same planes); 0 on success. */
typedef int (*FjoDiffusionFn)(const FjDiffusion* d, double* const planes[3], int W, int H, void* user);
/* Makes fjo_render use `fn` for its camera and enlarger diffusion stages
 * (NULL restores the direct convolution). */
void fjo_plan_set_diffusion(FjoPlan* plan, FjoDiffusionFn fn, void* user);
/* The differentiable print route's two departures from a render: the print
 * curves read at print_gamma * log E (1 restores the render), and a fixed
 * C/M/Y density increment (three width x height planes, NULL for none) added
 * where the grain stage adds its increment, on a record whose grain is off. */
void fjo_plan_set_fit(FjoPlan* plan, double print_gamma, const float* c, const float* m, const float* y);
/* Makes fjo_render copy the C/M/Y float densities entering its grain stage
 * into three width x height buffers (NULL disarms). */
void fjo_plan_capture_grain_input(FjoPlan* plan, float* c, float* m, float* y);
/* Sets the frame inputs of the scanner-glare stream: the frame token
 * (floor of the time in frames) and the source image origin (0 by default). */
void fjo_plan_set_frame(FjoPlan* plan, uint64_t frame_token, int origin_x, int origin_y);

/* Film TC-LUT build (rust reconstruction.rs:594-677): inputs as the
 * reference's FjFilmTcLutInput (tests/ffi/juicer_test_values.h:964-973) with
 * plain pointers; `out` receives 192 x 192 x 4 floats (R, G, B, 0). Returns 0
 * or one of the FJO_TC_* failures (the reference's TcError variants). */
typedef struct FjoTcLutInput {
    const float* spectra;              /* 192 x 192 x 81, (C, M, wavelength) */
    const float* sensitivity_rgb;      /* 81 x 3 */
    const float* reference_illuminant; /* 81 */
    float projection_white_xyz[3];
    float spectral_blur;
    uint32_t method;                   /* 0 Hanatos, 2 Arctic */
    uint32_t apply_surface;
    const float* surface_rgb;          /* 3 x 15 */
    uint32_t compression_active, hull_available;
    float hull_center_xy[2];
    const float* hull_xy;              /* 1025 x 2 */
} FjoTcLutInput;
enum {
    FJO_TC_INVALID_ILLUMINANT = 1,
    FJO_TC_INVALID_BLUR,
    FJO_TC_INVALID_KERNEL,
    FJO_TC_IMPOSSIBLE_KERNEL,
    FJO_TC_INVALID_ARCTIC_WHITE,
    FJO_TC_INVALID_NEUTRAL,
    FJO_TC_INVALID_SURFACE_WHITE,
    FJO_TC_NONFINITE_INTEGRATED,
    FJO_TC_NONFINITE_BLURRED,
    FJO_TC_MISSING_HULL,
    FJO_TC_UNUSABLE_COMPRESSED,
    FJO_TC_NONFINITE_REMAPPED,
    FJO_TC_UNUSABLE_SAMPLING,
};
int fjo_tc_lut_build(const FjoTcLutInput* input, float* out);
/* FNV-1a 64 of `bytes` bytes. */
uint64_t fjo_fnv1a64(const void* data, size_t bytes);

/* Philox4x32-10 (Salmon et al. 2011) of one 128-bit counter under a 64-bit key. */
void fjo_philox4x32_10(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]);
/* Scanner glare: the frame's stream seed, one pixel's glare percent, and the
 * unblurred field of a W x H window whose first pixel is (origin_x, origin_y)
 * (medium 1 for print routes, 0 for film). */
uint64_t fjo_glare_seed(uint64_t recipe_hash, uint64_t frame_token, int origin_x, int origin_y);
void fjo_glare_uniforms(uint64_t seed, int64_t abs_x, int64_t abs_y, uint32_t medium, float u[2]);
double fjo_glare_sample(uint64_t seed, int64_t abs_x, int64_t abs_y, uint32_t medium, float percent, float roughness);
void fjo_glare_plane(const FjScannerPostEffects* post, uint64_t seed, uint32_t medium, int origin_x, int origin_y,
                     int W, int H, double* out);

/* Builds host state for one prepared record (scanner LUT, curves). Returns
 * NULL and writes a message on unsupported input. */
FjoPlan* fjo_plan_create(const FjPreparedHostData* prepared, char* message, size_t capacity);
void fjo_plan_destroy(FjoPlan* plan);

/* Exposes the plan's scanner LUT (pointers stay valid until destroy). */
int fjo_plan_scanner_lut(const FjoPlan* plan, int* resolution, const float** log2_xyz,
                         const float** slope_c, const float** slope_m, const float** slope_y,
                         const float** cell_min, const float** cell_max);

/* Renders one full frame in double precision.
 *   src: host RGB/RGBA float32, row stride in floats.
 *   dst: host float32 output of the same component count (encoded, clamped),
 *        may be NULL.
 *   linear_rgb: optional [height][width][3] doubles: scanner-linear output
 *        after gamut compression and post effects, before output encoding.
 *   pixel_size_um / seed inputs come from the prepared record.
 * Returns 0 on success; *error_flags receives the scan/DIR error bits. */
/* Highlight boost of the camera carrier (three planes of n doubles, in
 * place; JuicerCudaHighlightBoost.cu). Returns 0 (applied or bypassed),
 * 6 for an invalid carrier, 7 when the peak exceeds FLT_MAX. */
int fjo_highlight_boost(const FjHighlightBoost* boost, double* const planes[3], size_t n);

/* Camera auto-exposure scale of a frame (JuicerCudaAutoExposure.cu): 1 when
 * the record's auto exposure is off, -1 when its source bounds are not the
 * rendered frame (0, 0, width, height). */
double fjo_auto_exposure(const FjoPlan* plan, const float* src, int width, int height, int components,
                         ptrdiff_t src_stride_floats);

int fjo_render(const FjoPlan* plan, const float* src, int width, int height, int components,
               ptrdiff_t src_stride_floats, float* dst, ptrdiff_t dst_stride_floats,
               double* linear_rgb, int threads, uint32_t* error_flags);

/* Per-pixel stages of the chain, the functions fjo_render composes, exposed
 * so that tests can drive each with the reference's own fixture words. */
/* Decode -> RGB->XYZ -> adapt -> Mallett (81-λ sum, mid-gray scale applied)
 * or TC-LUT exposure; xyz: working XYZ, srgb: clamped linear sRGB (Mallett
 * only), sums_bgr: the unclamped Mallett sums (Mallett only), raw_bgr: the
 * B/G/R raw exposure (JuicerCudaDeviceHelpers.cuh:509-703). */
void fjo_film_raw(const FjFilmExposure* exposure, const float rgb[3], double xyz[3], double srgb[3], double sums_bgr[3],
                  double raw_bgr[3]);
/* TC-LUT exposure of a working XYZ (tri2quad + Mitchell bicubic x brightness,
 * JuicerCudaDeviceHelpers.cuh:233-327, :563-617); raw in B/G/R order. */
void fjo_tc_exposure(const float* tc_lut_rgba, const double xyz[3], double raw_bgr[3]);
/* The 81-λ print integral of a C/M/Y density (print_spectral_integrate_device,
 * JuicerCudaDeviceHelpers.cuh:866-933), before normalizer and preflash. */
void fjo_print_integral(const FjPrint* print, const double density_cmy[3], double raw[3]);
/* Development curve lookup (JuicerCudaDeviceHelpers.cuh:34-89). */
double fjo_curve_at(const float* x, const float* y, int n, int stride, double log_exposure, double gamma);
/* Guarded Langmuir v (k + ref) / (k + v) (JuicerCudaDeviceHelpers.cuh:1073-1123). */
int fjo_langmuir(double v, double k, double ref, double* out);
/* Tricubic PCHIP lookup of the plan's scanner LUT at normalized C/M/Y
 * (JuicerCudaDeviceHelpers.cuh:407-506); 0 for non-finite input. */
int fjo_plan_scan_lut(const FjoPlan* plan, const double normalized_cmy[3], double log2_xyz[3]);
/* Scanner stage of one density: LUT, correction, CAT02, RGB, gamut. */
int fjo_plan_scan(const FjoPlan* plan, const double density_cmy[3], double rgb[3]);
/* OkLch output gamut compression (src/GamutCompression.h:178-213). */
void fjo_gamut_compress(const FjOutputColor* color, double rgb[3]);
/* Output matrix, CCTF and [0, 1] clamp (JuicerCudaScanPipeline.cu:99-190). */
void fjo_encode(const FjOutputColor* color, double rgb[3]);

#ifdef __cplusplus
}
#endif

#endif
