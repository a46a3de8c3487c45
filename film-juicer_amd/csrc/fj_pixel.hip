// fj_pixel.hip — fused per-pixel spectral chain for gfx950.
//
// One persistent launch walks the render window in 256-pixel tiles. Each block
// first stages the density curves and the compacted print spectra into LDS
// (binary searches and the 81-λ loop then hit LDS, not L2), then every lane
// runs the whole chain for one pixel in registers:
//   input decode -> XYZ -> TC-LUT Mitchell bicubic -> exposure -> film curves
//   -> DIR couplers -> [81-λ print exposure -> print curves] -> scanner PCHIP
//   LUT -> exp2 -> correction -> CAT02/XYZ->RGB -> OkLch gamut -> encode.
// Reference semantics: src/Cuda/JuicerCudaDeviceHelpers.cuh (a1-a6, a11-a15)
// and run_scan_pixel (src/Cuda/Scan/JuicerCudaScanPipeline.cu:1700-1859).
#include <hip/hip_runtime.h>

#include "fj_device.h"
#include "fj_pixel_math.h"

namespace fjx {

struct LdsLayout {
    float* film;        // 4n or 7n floats
    float* print;       // 4n' floats
    UniformF4 spec;     // 2k float4, scalar loads from the device table
    UniformF4 mallett;  // 2k float4, scalar loads from the device table
};

template <bool kPrint>
__device__ __forceinline__ LdsLayout stage_tables(const PixelParams& p, float* lds) {
    LdsLayout L;
    const int film_count = (p.dev.dir_mode ? 7 : 4) * p.dev.n;
    const int print_count = kPrint ? 4 * p.pr.n : 0;
    L.spec = uniform_table(p.pr.spec);
    L.mallett = uniform_table(p.ex.mallett);
    L.film = lds;
    L.print = L.film + film_count;
    for (int i = threadIdx.x; i < film_count; i += blockDim.x) L.film[i] = p.dev.curves[i];
    for (int i = threadIdx.x; i < print_count; i += blockDim.x) L.print[i] = p.pr.curves[i];
    __syncthreads();
    return L;
}

// Film exposure of one pixel: from the camera carrier with the route
// correction (JuicerCudaDeviceHelpers.cuh:746-755, :838-848) when a carrier
// exists, else straight from the source pixel.
__device__ __forceinline__ void film_exposure(const PixelParams& p, UniformF4 mallett, const float in[4],
                                              long long idx, float E[3]) {
    if (p.carrier[0]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) E[c] = fmaxf(0.0f, p.carrier[c][idx] * p.ex.route_scale);
        return;
    }
    camera_exposure(p.ex, mallett, in, E);
}

// What a development stage writes: the full chain to the frame (or linear
// planes), the density planes only (grain follows), or the print-linear
// planes only (enlarger diffusion follows). Fixed per launch, so a split
// stage is compiled without the code of the stages after it.
enum PixelOut : int { kOutFull = 0, kOutDensity = 1, kOutPrintLinear = 2 };

template <bool kPrint, bool kLds, int kMode, int kOut>
__device__ __forceinline__ void pixel_stage_body(const PixelParams& p, int tiles) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    LdsLayout L;
    if constexpr (kLds) {
        L = stage_tables<kPrint>(p, lds);
    } else {
        L.film = const_cast<float*>(p.dev.curves);
        L.print = const_cast<float*>(p.pr.curves);
        L.spec = uniform_table(p.pr.spec);
        L.mallett = uniform_table(p.ex.mallett);
    }
    const int width = p.io.width;
    const long long total = static_cast<long long>(width) * p.io.height;
    unsigned int errors = 0u;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long long idx = static_cast<long long>(tile) * blockDim.x + threadIdx.x;
        if (idx >= total) continue;
        const int y = static_cast<int>(idx / width);
        const int x = static_cast<int>(idx - static_cast<long long>(y) * width);
        float in[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        if constexpr (kMode != kDirFinal && kMode != kEnlargerFinal && kMode != kDensityFinal) {
            const float* sp = reinterpret_cast<const float*>(p.io.src + y * p.io.src_row_bytes) + x * p.io.comps;
            if (p.io.comps == 4) {
                const float4 v = *reinterpret_cast<const float4*>(sp);
                in[0] = v.x; in[1] = v.y; in[2] = v.z; in[3] = v.w;
            } else {
                in[0] = sp[0]; in[1] = sp[1]; in[2] = sp[2];
            }
        }
        float D[3];
        if constexpr (kMode == kDensityFinal) {
#pragma unroll
            for (int c = 0; c < 3; ++c) D[c] = p.dens[c][idx];
        } else if constexpr (kMode == kEnlargerFinal) {
            float lin[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) lin[c] = p.plin[c][idx];
            print_develop_linear(p.pr, L.print, lin, D);
        } else if constexpr (kMode == kCarrier) {
            // Camera carrier: exposure with the manual scale only
            // (camera_film_linear_exposure_kernel, JuicerCudaFilmPipeline.cu:3103-3146).
            float E[3];
            camera_exposure(p.ex, L.mallett, in, E);
#pragma unroll
            for (int c = 0; c < 3; ++c) p.carrier[c][idx] = E[c];
            continue;
        } else if constexpr (kMode == kDirSource) {
            float E[3], corr[3], log_raw[3];
            film_exposure(p, L.mallett, in, idx, E);
            errors |= dir_source(p.dev, L.film, E, corr, log_raw);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                p.dir_planes[c][idx] = corr[c];
                p.dir_planes[3 + c][idx] = log_raw[c];
            }
            continue;
        } else if constexpr (kMode == kDirFinal) {
            float filtered[3], log_raw[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                filtered[c] = p.dir_planes[c][idx];
                log_raw[c] = p.dir_planes[3 + c][idx];
            }
            errors |= develop_dir_final(p.dev, L.film, log_raw, filtered, D);
        } else {
            float E[3];
            film_exposure(p, L.mallett, in, idx, E);
            errors |= develop_film(p.dev, L.film, E, D);
        }
        if constexpr (kOut == kOutDensity) {
#pragma unroll
            for (int c = 0; c < 3; ++c) p.dens[c][idx] = D[c];
            continue;
        } else if constexpr (kOut == kOutPrintLinear) {
            float lin[3];
            print_linear(p.pr, L.spec, D, lin);
#pragma unroll
            for (int c = 0; c < 3; ++c) p.plin[c][idx] = lin[c];
            continue;
        } else if constexpr (kPrint && kMode != kEnlargerFinal) {
            print_develop(p.pr, L.spec, L.print, D);
        }
        float rgb[3];
        const bool ok = scan_colour(p.sc, p.co, D, p.sc.glare ? p.sc.glare[idx] : 0.0f, rgb);
        if (!ok) {
            errors |= 1u;
            rgb[0] = rgb[1] = rgb[2] = 0.0f;
        }
        if (p.linear_out) {
            const long long plane = total;
            p.linear_planes[idx] = rgb[0];
            p.linear_planes[plane + idx] = rgb[1];
            p.linear_planes[2 * plane + idx] = rgb[2];
            continue;
        }
        if constexpr (kMode == kDirFinal || kMode == kEnlargerFinal || kMode == kDensityFinal) {
            const float* sp = reinterpret_cast<const float*>(p.io.src + y * p.io.src_row_bytes) + x * p.io.comps;
            if (p.io.comps == 4) in[3] = sp[3];
        }
        if (ok) encode_output(p.co, rgb);
        float* dp = reinterpret_cast<float*>(p.io.dst + y * p.io.dst_row_bytes) + x * p.io.comps;
        if (p.io.comps == 4) {
            *reinterpret_cast<float4*>(dp) = make_float4(rgb[0], rgb[1], rgb[2], in[3]);
        } else {
            dp[0] = rgb[0]; dp[1] = rgb[1]; dp[2] = rgb[2];
        }
    }
    if (errors && p.error_word) atomicOr(p.error_word, errors);
}

template <bool kPrint, bool kLds, int kMode, int kOut = kOutFull>
__global__ __launch_bounds__(256) void pixel_stage_kernel(PixelParams p, int tiles) {
    pixel_stage_body<kPrint, kLds, kMode, kOut>(p, tiles);
}

// The print + scan stage after enlarger diffusion: capped at five waves per
// SIMD it fits 96 VGPRs without scratch (from three waves at 129).
template <>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void
pixel_stage_kernel<true, true, kEnlargerFinal, kOutFull>(PixelParams p, int tiles) {
    pixel_stage_body<true, true, kEnlargerFinal, kOutFull>(p, tiles);
}
template <>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void
pixel_stage_kernel<true, false, kEnlargerFinal, kOutFull>(PixelParams p, int tiles) {
    pixel_stage_body<true, false, kEnlargerFinal, kOutFull>(p, tiles);
}

__global__ void clear_word_kernel(unsigned int* word) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *word = 0u;
}

}  // namespace fjx

// ---------------------------------------------------------------------------
// launchers (C++ linkage inside the library)
// ---------------------------------------------------------------------------
namespace fjx {

size_t fused_lds_bytes(const PixelParams& p) {
    const size_t film = static_cast<size_t>(p.dev.dir_mode ? 7 : 4) * p.dev.n;
    const size_t prn = p.print ? static_cast<size_t>(4) * p.pr.n : 0;
    return (film + prn) * sizeof(float);
}

hipError_t launch_clear_word(unsigned int* word, hipStream_t stream) {
    hipLaunchKernelGGL(clear_word_kernel, dim3(1), dim3(64), 0, stream, word);
    return hipGetLastError();
}

// Output kind of a development stage (the order of the checks is the
// chain's: densities for grain first, then print-linear planes).
static int out_kind(const PixelParams& p, int mode) {
    if ((mode == kFused || mode == kDirFinal) && p.dens[0]) return kOutDensity;
    if (p.print && mode != kEnlargerFinal && p.plin[0]) return kOutPrintLinear;
    return kOutFull;
}

// Launches a persistent pixel kernel with four times the blocks the GPU
// holds at once (blocks per CU for this kernel and LDS size x CUs, at most one
// per tile): whole rounds of resident blocks, and enough of them that blocks
// finishing early pick up work (measured at UHD: 1x, 2x, 4x, 8x and one block
// per tile give 294, 301, 302, 299 and 293 frames/s).
constexpr int kPixelGridRounds = 4;

template <typename Kernel>
static void launch_resident(Kernel kernel, const PixelParams& p, int cus, int tiles, size_t lds, hipStream_t stream) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const long long resident = static_cast<long long>(cus) * per_cu * kPixelGridRounds;
    const int grid = tiles < resident ? tiles : static_cast<int>(resident);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), lds, stream, p, tiles);
}

template <int kMode, int kOut>
static void launch_out(const PixelParams& p, int cus, int tiles, size_t lds, bool use_lds, hipStream_t stream) {
    if (p.print) {
        if (use_lds)
            launch_resident(pixel_stage_kernel<true, true, kMode, kOut>, p, cus, tiles, lds, stream);
        else
            launch_resident(pixel_stage_kernel<true, false, kMode, kOut>, p, cus, tiles, 0, stream);
    } else {
        if (use_lds)
            launch_resident(pixel_stage_kernel<false, true, kMode, kOut>, p, cus, tiles, lds, stream);
        else
            launch_resident(pixel_stage_kernel<false, false, kMode, kOut>, p, cus, tiles, 0, stream);
    }
}

template <int kMode>
static hipError_t launch_mode(const PixelParams& p, int cus, int tiles, size_t lds, bool use_lds, hipStream_t stream) {
    switch (out_kind(p, kMode)) {
        case kOutDensity:
            if constexpr (kMode == kFused || kMode == kDirFinal)
                launch_out<kMode, kOutDensity>(p, cus, tiles, lds, use_lds, stream);
            break;
        case kOutPrintLinear:
            if constexpr (kMode != kEnlargerFinal)
                launch_out<kMode, kOutPrintLinear>(p, cus, tiles, lds, use_lds, stream);
            break;
        default: launch_out<kMode, kOutFull>(p, cus, tiles, lds, use_lds, stream); break;
    }
    return hipGetLastError();
}

hipError_t launch_pixel_stage(const PixelParams& p, int mode, int compute_units, hipStream_t stream) {
    const long long total = static_cast<long long>(p.io.width) * p.io.height;
    if (total <= 0) return hipSuccess;
    const int block = 256;
    const long long tiles_ll = (total + block - 1) / block;
    if (tiles_ll > 0x7fffffffLL) return hipErrorInvalidValue;
    const int tiles = static_cast<int>(tiles_ll);
    const size_t lds = fused_lds_bytes(p);
    const bool use_lds = lds <= 64 * 1024 && p.dev.n <= kMaxLdsCurve && (!p.print || p.pr.n <= kMaxLdsCurve);
    const int cus = compute_units > 0 ? compute_units : 1;
    switch (mode) {
        case kDirSource: {
            // The source stage needs neither print nor scanner tables.
            PixelParams q = p;
            q.print = 0;
            const size_t lq = fused_lds_bytes(q);
            if (use_lds)
                launch_resident(pixel_stage_kernel<false, true, kDirSource>, q, cus, tiles, lq, stream);
            else
                launch_resident(pixel_stage_kernel<false, false, kDirSource>, q, cus, tiles, 0, stream);
            return hipGetLastError();
        }
        case kCarrier: {
            // Exposure only, with the camera scale (no route correction).
            PixelParams q = p;
            q.print = 0;
            q.ex.scale = p.ex.camera_scale;
            q.ex.with_route = 0;
            const size_t lq = fused_lds_bytes(q);
            if (use_lds)
                launch_resident(pixel_stage_kernel<false, true, kCarrier>, q, cus, tiles, lq, stream);
            else
                launch_resident(pixel_stage_kernel<false, false, kCarrier>, q, cus, tiles, 0, stream);
            return hipGetLastError();
        }
        case kDirFinal: return launch_mode<kDirFinal>(p, cus, tiles, lds, use_lds, stream);
        case kDensityFinal: {
            // Print exposure (or its linear planes), scanner and colour: no
            // exposure or film tables.
            PixelParams q = p;
            q.ex.mallett = nullptr;
            q.dev.n = 0;
            q.dev.dir_mode = 0;
            const size_t lq = fused_lds_bytes(q);
            const bool lds_ok = lq <= 64 * 1024 && (!q.print || q.pr.n <= kMaxLdsCurve);
            return launch_mode<kDensityFinal>(q, cus, tiles, lq, lds_ok, stream);
        }
        case kEnlargerFinal: {
            // Print curves, scanner and colour only: no film tables.
            if (!p.print) return hipErrorInvalidValue;
            PixelParams q = p;
            q.ex.mallett = nullptr;
            q.dev.n = 0;
            q.pr.k = 0;
            const size_t lq = fused_lds_bytes(q);
            const bool lds_ok = lq <= 64 * 1024 && q.pr.n <= kMaxLdsCurve;
            if (lds_ok)
                launch_resident(pixel_stage_kernel<true, true, kEnlargerFinal>, q, cus, tiles, lq, stream);
            else
                launch_resident(pixel_stage_kernel<true, false, kEnlargerFinal>, q, cus, tiles, 0, stream);
            return hipGetLastError();
        }
        default: return launch_mode<kFused>(p, cus, tiles, lds, use_lds, stream);
    }
}

}  // namespace fjx
