// fj_device.h — device-visible parameter blocks of the MI355X render path.
//
// Everything here is plain data passed by value as kernel arguments; tables
// live in device memory (HBM, L2/MALL-resident) and are staged into LDS by the
// kernels that search them. Channel order on device follows the reference:
// film layers B/G/R, dyes C/M/Y (JuicerCudaFilmPayloads.cpp:157-190).
#pragma once

#include <stdint.h>

namespace fjx {

constexpr int kTcExtent = 192;     // TC LUT side (reference kFilmTcLutExtent)
constexpr int kSpectralSamples = 81;
constexpr int kGamutL = 64;
constexpr int kGamutH = 720;
constexpr int kMaxLdsCurve = 512;  // curves longer than this are read from global memory

// Input decode + film exposure.
struct ExposureParams {
    const float4* tc_lut;   // [192*192] (R, G, B, pad); null for Mallett
    float rgb_to_xyz[9];    // input_xyz_adapt * input_rgb_to_xyz when adaptation is on
    int decode_cctf;        // 0, or the FJ_INPUT_* space to decode
    float scale;            // exp2(EV) * route correction (validated > 0)
    float camera_scale;     // exp2(EV) alone: the camera carrier's scale
    float route_scale;      // route correction applied when a carrier is developed
    // Auto exposure: the metered scale (device scalar) replaces `scale` by
    // (ae * camera_scale) [* route_scale when with_route], validated > 0.
    const float* ae;
    int with_route;
    // Mallett 2019 reconstruction: k pairs {basis_r*illum, basis_g*illum,
    // basis_b*illum, 0}, {sens_B, sens_G, sens_R, 0}; rgb_to_srgb maps the
    // input to linear sRGB; green mid-gray scale applied to the result.
    const float4* mallett;
    int mallett_k;
    float mallett_scale;
    float rgb_to_srgb[9];
};

// Film development: [x | yB | yG | yR | cxB | cxG | cxR] each n floats.
struct DevelopParams {
    const float* curves;
    int n;
    float gamma_bgr[3];
    int dir_mode;           // FJ_DIR_*
    float M[9];             // [donor BGR][receiver BGR]
    float dmax[3], dref[3], donor_k[3], recv_cref[3], recv_kr[3];
};

// Print exposure + development.
// spec: k pairs of float4 {e_c, e_m, e_y, e_base} (already scaled by -log2 10)
// and {s_c*illum, s_m*illum, s_y*illum, 0}. curves: [x | yC | yM | yY].
struct PrintParams {
    const float4* spec;
    int k;
    float normalizer;
    float preflash[3];      // preflash_raw * preflash_exposure (0 when inactive)
    float outer_scale;      // print exposure * scanner exposure scale
    const float* curves;
    int n;
};

// Scanner LUT (interleaved PCHIP nodes) and colour.
// nodes: 3 float4 per node {v0 v1 v2 sc0}{sc1 sc2 sm0 sm1}{sm2 sy0 sy1 sy2}
// cells: 2 float4 per cell {min0 min1 min2 max0}{max1 max2 - -}
struct ScanParams {
    const float4* nodes;
    const float4* cells;
    int res;
    float dmin[3];
    float inv_span[3];
    int correction;
    float corr_slope, corr_offset;
    float xyz_to_rgb[9];    // xyz_to_rgb * cat02
    const float* glare;     // scanner glare percent per pixel, or null
    float glare_xyz[3];     // 0.01 x the scan illuminant's XYZ
};

struct ColorParams {
    int gamut;
    const float* cmax;      // [64][720]
    float native_to_xyz[9];
    float xyz_to_native[9];
    float lknee[3];
    float cknee[3];
    int encode;             // apply CCTF
    int matrix;             // apply dwg_to_output
    float to_output[9];
    int cctf;
    float gamma, a, b, c, d, cutoff;
};

// A frame window: pixels [x0, x0+w) x [y0, y0+h) in absolute coordinates.
struct FrameIO {
    const char* src;        // address of pixel (x0, y0)
    long long src_row_bytes;
    char* dst;
    long long dst_row_bytes;
    int width, height, comps;
};

struct PixelParams {
    FrameIO io;
    ExposureParams ex;
    DevelopParams dev;
    PrintParams pr;
    ScanParams sc;
    ColorParams co;
    unsigned int* error_word;  // bit 0 scan error; bits 8.. DIR failure
    int print;
    int linear_out;            // write linear planes instead of encoded RGBA
    float* linear_planes;      // 3 planes of width*height when linear_out
    // Spatial-DIR planes (width*height each): [0..2] corrections Y/M/C
    // (raw from the source stage, filtered into the final stage), [3..5] log
    // raw exposure B/G/R.
    float* dir_planes[6];
    // Camera carrier (film linear exposure B/G/R, width*height each). When
    // set, the exposure of the source/fused stages is read from it instead
    // of computed from the source image; the carrier stage writes it.
    float* carrier[3];
    // Enlarger print-linear exposure (C/M/Y-forming, width*height each). When
    // set, the fused and DIR-final stages stop at the print exposure and
    // write it here; the enlarger-final stage develops the print from it.
    float* plin[3];
    // Film density C/M/Y planes (width*height each) for the visual-grain
    // stage: the fused and DIR-final stages stop after development and write
    // them; the density-final stage continues from them.
    float* dens[3];
};

// Pixel stage of one launch.
enum PixelMode : int { kFused = 0, kDirSource = 1, kDirFinal = 2, kCarrier = 3, kEnlargerFinal = 4, kDensityFinal = 5 };

// Post stage: unsharp mask on the three linear planes, output encoding,
// alpha copied from the source (JuicerCudaScanPipeline.cu:2733-2878).
struct PostParams {
    FrameIO io;
    ColorParams co;
    const float* planes;       // 3 planes width*height, R/G/B
    const float* weights;      // 2r+1 unsharp weights
    int radius;                // 0: no unsharp
    float amount;
    // Gate weave: the output pixel (x, y) reads the planes at
    // R (x - c, y - c) + c + (dx, dy) through a Mitchell filter; c is the
    // window centre. Set only with radius 0 (the unsharp is materialized).
    int weave;
    float weave_dx, weave_dy, weave_cos, weave_sin;
};

}  // namespace fjx
