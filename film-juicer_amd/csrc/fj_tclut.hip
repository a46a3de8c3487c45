// fj_tclut.hip — the film TC-LUT build's table passes on the GPU.
//
// The TC LUT (rust/film-juicer-core/src/reconstruction.rs:594-677) is an FP32
// table whose words are the contract, so the passes here keep the
// reference's FP32 operations in its order, without contraction:
//   * the optional spectral blur along wavelength (:540-569): one thread per
//     (cell, wavelength) sample, the taps summed in kernel order;
//   * the sensitivity integration (:627-650): one thread per cell, the 81
//     samples summed in order, divided by the neutral response, times the
//     cell's surface factor when given;
//   * the input-gamut remap (gamut.rs:256-364): one thread per cell, the hull
//     compression of the cell's chromaticity in FP64 (ray against the 1024
//     hull edges, knee) and a bilinear lookup of the integrated table.
// Scalar preparation with transcendental functions (blur weights, surface
// factors, the neutral response, the remap's sqrt grid) runs on the host
// with the same library functions as the reference.
#include "fj_tclut.h"

#include <cmath>

#pragma clang fp contract(off)

namespace fjx {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int wavelength_reflect(int i) {
    int v = i % 162;
    if (v < 0) v += 162;
    return v >= 81 ? 161 - v : v;
}

__device__ __forceinline__ float clamp01(float v) { return isnan(v) ? v : (v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v)); }

__global__ __launch_bounds__(kThreads) void tc_blur_kernel(TcBlurArgs a) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= kTcCells * kTcSamples) return;
    const int cell = i / kTcSamples, s = i - cell * kTcSamples;
    const float* src = a.src + static_cast<size_t>(cell) * kTcSamples;
    float v = 0.0f;
    for (int k = 0; k <= 2 * a.radius; ++k) v += a.kernel[k] * src[wavelength_reflect(s + k - a.radius)];
    a.dst[i] = v;
    if (!isfinite(v)) atomicMin(a.first_failure, static_cast<unsigned int>(i));
}

__global__ __launch_bounds__(kThreads) void tc_integrate_kernel(TcIntegrateArgs a) {
    __shared__ float sens[kTcSamples * 3];
    __shared__ float ill[kTcSamples];
    for (int i = threadIdx.x; i < kTcSamples * 3; i += kThreads) sens[i] = a.sensitivity[i];
    for (int i = threadIdx.x; i < kTcSamples; i += kThreads) ill[i] = a.illuminant[i];
    __syncthreads();
    const int cell = blockIdx.x * kThreads + threadIdx.x;
    if (cell >= kTcCells) return;
    const float* sp = a.spectra + static_cast<size_t>(cell) * kTcSamples;
    float raw[3] = {0.0f, 0.0f, 0.0f};
    for (int s = 0; s < kTcSamples; ++s) {
        float energy = sp[s];
        if (a.relight) energy = energy * ill[s];
#pragma unroll
        for (int c = 0; c < 3; ++c) raw[c] = raw[c] + energy * sens[s * 3 + c];
    }
    float* out = a.out + static_cast<size_t>(cell) * 4;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float value = raw[c] / a.denominator[c];
        if (a.surface_factor) value = value * a.surface_factor[static_cast<size_t>(cell) * 3 + c];
        out[c] = value;
        if (!isfinite(value)) atomicMin(a.first_failure, static_cast<unsigned int>(cell) * 8u + static_cast<unsigned>(c));
    }
    out[3] = 0.0f;
}

// Distance from the hull centre along a unit direction to the nearest edge
// crossing ahead of it (gamut.rs:256-286).
__device__ double ray_polygon_distance(double c0, double c1, double d0, double d1, const float* xy) {
    double minimum = INFINITY;
    for (int e = 0; e < 1024; ++e) {
        const double ax = xy[2 * e], ay = xy[2 * e + 1];
        const double ex = static_cast<double>(xy[2 * e + 2]) - ax, ey = static_cast<double>(xy[2 * e + 3]) - ay;
        const double den = d0 * ey - d1 * ex;
        if (fabs(den) <= 1e-12) continue;
        const double ox = c0 - ax, oy = c1 - ay;
        const double distance = (-ox * ey + oy * ex) / den;
        if (distance > 1e-9 && distance < minimum) {
            const double position = (-ox * d1 + oy * d0) / den;
            if (position >= -1e-9 && position <= 1.0 + 1e-9) minimum = distance;
        }
    }
    return minimum;
}

__global__ __launch_bounds__(kThreads) void tc_remap_kernel(TcRemapArgs a) {
    __shared__ float hull[2 * 1025];
    for (int i = threadIdx.x; i < 2 * 1025; i += kThreads) hull[i] = a.hull_xy[i];
    __syncthreads();
    const int cell = blockIdx.x * kThreads + threadIdx.x;
    if (cell >= kTcCells) return;
    const int c = cell / kTcSide, m = cell - c * kTcSide;
    const float root = a.roots[c];
    float xy[2] = {1.0f - root, a.steps[m] * root};
    // compress (gamut.rs:296-321)
    const double c0 = a.hull_center[0], c1 = a.hull_center[1];
    const double dx = static_cast<double>(xy[0]) - c0, dy = static_cast<double>(xy[1]) - c1;
    const double distance = hypot(dx, dy);
    if (distance >= 1e-9) {
        const double d0 = dx / distance, d1 = dy / distance;
        const double boundary = ray_polygon_distance(c0, c1, d0, d1, hull);
        if (!isfinite(boundary) || boundary <= 1e-12) {
            atomicMin(a.first_failure, static_cast<unsigned int>(cell) * 8u);
            return;
        }
        double k = distance / boundary;
        if (k > 0.815) {
            const double scale = 1.0 - 0.815;
            const double n = (k - 0.815) / scale;
            k = 0.815 + scale * (n / pow(1.0 + pow(n, 1.2), 1.0 / 1.2));
        }
        const double cd = k * boundary;
        const double r0 = c0 + d0 * cd, r1 = c1 + d1 * cd;
        if (!isfinite(r0) || !isfinite(r1)) {
            atomicMin(a.first_failure, static_cast<unsigned int>(cell) * 8u);
            return;
        }
        xy[0] = static_cast<float>(r0);
        xy[1] = static_cast<float>(r1);
    }
    // tri2quad (reconstruction.rs:352-356)
    const float x = 1.0f - xy[0];
    const float den = x < 1e-10f ? 1e-10f : x;
    const float tc0 = clamp01(x * x), tc1 = clamp01(xy[1] / den);
    // bilinear (gamut.rs:322-341)
    const float cp = clamp01(tc0) * 191.0f, mp = clamp01(tc1) * 191.0f;
    float* out = a.out + static_cast<size_t>(cell) * 4;
    out[3] = 0.0f;
    if (!isfinite(cp) || !isfinite(mp)) {
        atomicMin(a.first_failure, static_cast<unsigned int>(cell) * 8u + 1u);
        return;
    }
    const int ca = static_cast<int>(floorf(cp)), ma = static_cast<int>(floorf(mp));
    const int cb = ca + 1 < 191 ? ca + 1 : 191, mb = ma + 1 < 191 ? ma + 1 : 191;
    const float cf = cp - static_cast<float>(ca), mf = mp - static_cast<float>(ma);
    const float* t = a.integrated;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float v00 = t[(static_cast<size_t>(ca) * kTcSide + ma) * 4 + ch];
        const float v10 = t[(static_cast<size_t>(cb) * kTcSide + ma) * 4 + ch];
        const float v01 = t[(static_cast<size_t>(ca) * kTcSide + mb) * 4 + ch];
        const float v11 = t[(static_cast<size_t>(cb) * kTcSide + mb) * 4 + ch];
        const float row0 = v00 + cf * (v10 - v00);
        const float row1 = v01 + cf * (v11 - v01);
        const float v = row0 + mf * (row1 - row0);
        out[ch] = v;
        if (!isfinite(v)) atomicMin(a.first_failure, static_cast<unsigned int>(cell) * 8u + 2u + static_cast<unsigned>(ch));
    }
}

}  // namespace

hipError_t launch_tc_blur(const TcBlurArgs& a, hipStream_t s) {
    const int n = kTcCells * kTcSamples;
    hipLaunchKernelGGL(tc_blur_kernel, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tc_integrate(const TcIntegrateArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(tc_integrate_kernel, dim3((kTcCells + kThreads - 1) / kThreads), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tc_remap(const TcRemapArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(tc_remap_kernel, dim3((kTcCells + kThreads - 1) / kThreads), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace fjx
