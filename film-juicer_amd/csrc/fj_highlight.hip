// fj_highlight.hip — highlight boost of the camera carrier for gfx950.
//
// Reference operator: src/Cuda/Film/JuicerCudaHighlightBoost.cu (maximum and
// validity reduction :40-60, disposition and shape :71-111, per-sample boost
// :112-186). Over the three carrier planes: m = the largest sample; any
// non-finite or negative sample makes the carrier invalid. With protection p,
// sharpness k0 and peak gain g, samples s > p become
//   s + m (g - 1) f^2 N(k f) / N(k),  f = (s - p) / (m - p), k = k0 (m - p) / m,
//   N(t) = (e^t - t - 1) / t^2,
// so s = m maps to the peak m g and the curve leaves p with zero slope. The
// reference evaluates this in FP32 with carried residual words; here each
// boosted sample is evaluated in FP64 (only samples above p are touched).
// Dispositions: invalid carrier and a peak beyond FLT_MAX fail the frame
// (error-word bits), m <= p bypasses.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "fj_highlight.h"

namespace fjx {

namespace {

constexpr int kThreads = 256;

// g(t) = (e^t - t - 1) / t^2 for t >= 0, in FP64 (the per-frame constants).
__device__ __forceinline__ double normalized_exp(double t) {
    // By its Taylor series near zero.
    if (t >= 0.03125) return (expm1(t) - t) / (t * t);
    return 0.5 + t * (1.0 / 6.0 + t * (1.0 / 24.0 + t * (1.0 / 120.0 + t * (1.0 / 720.0 +
                 t * (1.0 / 5040.0 + t * (1.0 / 40320.0 + t * (1.0 / 362880.0 + t / 3628800.0)))))));
}

// g(t) in FP32 per pixel for t = t_hi + t_lo (|t_lo| a rounding residual):
// the series sum_k t^k / (k + 2)! below t = 1 (twelve terms: the remainder is
// under 1e-9 of the sum), else (e^t - t - 1) / t^2 with e^t = e^t_hi (1 + t_lo)
// to first order, so that the rounding of t (up to 1e-6 of e^t at t = 28)
// does not reach the result; the cancellation costs at most a factor 4.4 of
// expm1f's rounding.
__device__ __forceinline__ float normalized_expf(float t_hi, float t_lo) {
    if (t_hi >= 1.0f) {
        const float num = fmaf(expf(t_hi), t_lo, expm1f(t_hi) - t_hi) - t_lo;
        return num / (t_hi * fmaf(2.0f, t_lo, t_hi));
    }
    const float t = t_hi + t_lo;
    constexpr float c[12] = {1.0f / 2,        1.0f / 6,         1.0f / 24,        1.0f / 120,
                             1.0f / 720,      1.0f / 5040,      1.0f / 40320,     1.0f / 362880,
                             1.0f / 3628800,  1.0f / 39916800,  1.0f / 479001600, 1.0f / 6227020800.0f};
    float g = c[11];
#pragma unroll
    for (int k = 10; k >= 0; --k) g = fmaf(g, t, c[k]);
    return g;
}

__global__ __launch_bounds__(kThreads) void hb_reduce_kernel(HbArgs a) {
    float mx = 0.0f;
    unsigned int invalid = 0u;
    const long long step = static_cast<long long>(gridDim.x) * kThreads;
    for (long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; i < a.n; i += step) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = a.planes[c][i];
            if (isfinite(s) && s >= 0.0f)
                mx = fmaxf(mx, s);
            else
                invalid = 1u;
        }
    }
    __shared__ float s_m[kThreads];
    __shared__ unsigned int s_i[kThreads];
    s_m[threadIdx.x] = mx;
    s_i[threadIdx.x] = invalid;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_m[threadIdx.x] = fmaxf(s_m[threadIdx.x], s_m[threadIdx.x + o]);
            s_i[threadIdx.x] |= s_i[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.state->partial_max[blockIdx.x] = s_m[0];
        a.state->partial_invalid[blockIdx.x] = s_i[0];
    }
}

__global__ __launch_bounds__(kThreads) void hb_finalize_kernel(HbArgs a, int blocks) {
    __shared__ float s_m[kThreads];
    __shared__ unsigned int s_i[kThreads];
    float mx = 0.0f;
    unsigned int invalid = 0u;
    for (int b = threadIdx.x; b < blocks; b += kThreads) {
        mx = fmaxf(mx, a.state->partial_max[b]);
        invalid |= a.state->partial_invalid[b];
    }
    s_m[threadIdx.x] = mx;
    s_i[threadIdx.x] = invalid;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_m[threadIdx.x] = fmaxf(s_m[threadIdx.x], s_m[threadIdx.x + o]);
            s_i[threadIdx.x] |= s_i[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    HbState& st = *a.state;
    const double m = s_m[0];
    st.maximum = s_m[0];
    st.applied = 0;
    if (s_i[0]) {
        atomicOr(a.error_word, kHbInvalidCarrierBit);
    } else if (m <= a.protection) {
        // bypass
    } else if (m * a.peak_gain > static_cast<double>(FLT_MAX)) {
        atomicOr(a.error_word, kHbPeakOverflowBit);
    } else {
        const double range = m - a.protection;
        const double shape = a.sharpness * range / m;
        const float p_hi = static_cast<float>(a.protection);
        st.protection_hi = p_hi;
        st.protection_lo = static_cast<float>(a.protection - static_cast<double>(p_hi));
        st.inv_range = static_cast<float>(1.0 / range);
        const double slope = shape / range;
        st.slope_hi = static_cast<float>(slope);
        st.slope_lo = static_cast<float>(slope - static_cast<double>(st.slope_hi));
        st.amplitude = static_cast<float>(m * (a.peak_gain - 1.0) / normalized_exp(shape));
        st.peak = static_cast<float>(m * a.peak_gain);
        st.applied = 1;
    }
}

__global__ __launch_bounds__(kThreads) void hb_apply_kernel(HbArgs a) {
    const HbState& st = *a.state;
    if (!st.applied) return;
    const long long step = static_cast<long long>(gridDim.x) * kThreads;
    for (long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; i < a.n; i += step) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = a.planes[c][i];
            // s > p exactly, with p = hi + lo: at s == hi the sign of lo decides.
            const bool above = st.protection_lo < 0.0f ? s >= st.protection_hi : s > st.protection_hi;
            if (!above) continue;
            if (s == st.maximum) {
                a.planes[c][i] = st.peak;
                continue;
            }
            // d = s - p as a float pair (exact difference of s and hi, then lo),
            // t = d x slope as a float pair, f = d / range.
            const float d1 = s - st.protection_hi;
            const float bb = d1 - s;
            const float d1_err = (s - (d1 - bb)) + (-st.protection_hi - bb);
            const float d_hi = d1, d_lo = d1_err - st.protection_lo;
            const float t_hi = d_hi * st.slope_hi;
            const float t_lo = fmaf(d_hi, st.slope_hi, -t_hi) + fmaf(d_hi, st.slope_lo, d_lo * st.slope_hi);
            const float f = (d_hi + d_lo) * st.inv_range;
            a.planes[c][i] = fmaf(st.amplitude * f * f, normalized_expf(t_hi, t_lo), s);
        }
    }
}

}  // namespace

hipError_t launch_highlight_boost(const HbArgs& a, hipStream_t stream) {
    if (a.n <= 0) return hipSuccess;
    const long long want = (a.n + kThreads - 1) / kThreads;
    const int blocks = static_cast<int>(want < kHbMaxBlocks ? want : kHbMaxBlocks);
    hipLaunchKernelGGL(hb_reduce_kernel, dim3(blocks), dim3(kThreads), 0, stream, a);
    hipLaunchKernelGGL(hb_finalize_kernel, dim3(1), dim3(kThreads), 0, stream, a, blocks);
    hipLaunchKernelGGL(hb_apply_kernel, dim3(blocks), dim3(kThreads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace fjx
