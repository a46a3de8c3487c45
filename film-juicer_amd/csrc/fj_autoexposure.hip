// fj_autoexposure.hip — camera auto-exposure metering for gfx950.
//
// Reference operator: src/Cuda/JuicerCudaAutoExposure.cu (preview sampling
// :75-85, metering weights :87-153, weighted sums :314-407, finalisation
// :462-522, median histogram :524-660, launcher :663-908). The meter reads a
// preview grid (long edge <= 256) of the meter window, reconstructs each
// sample's film raw exposure (the camera exposure without any scale), and
// turns a weighted mean (or the histogram median) Y of the raw means into the
// device scale 2^(-log2 Y), which the exposure kernels multiply in.
//
// Here: one launch writes FP64 per-block partial sums (four lanes of sum Y
// and sum W) for the weighted methods, or per-block maxima for the median;
// a one-block launch reduces the partials in block order (deterministic) and
// writes the scale. The median adds a histogram launch (2048 bins, LDS then
// global integer atomics) and its own one-block finalisation.
#include <hip/hip_runtime.h>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_autoexposure.h"
#include "fj_pixel_math.h"

// The meter's decisions (ring and spot membership on the FP32 radius, the
// median's bin) are made on unfused single-precision operations.
#pragma clang fp contract(off)

namespace fjx {

namespace {

constexpr int kThreads = 256;
constexpr int kLanes = 4;

// Source coordinate of preview index i on an axis of `source` samples
// (preview_source_coordinate, JuicerCudaAutoExposure.cu:75-85).
__device__ __forceinline__ int preview_coord(int i, int preview, int source) {
    if (preview <= 0 || source <= 0) return 0;
    const long long num = (2LL * i + 1LL) * static_cast<long long>(source);
    const int c = static_cast<int>(num / (2LL * preview));
    return min(source - 1, max(0, c));
}

// Raw mean of preview sample (x, y); false when the sample lies outside the
// source image or its mean is not finite.
__device__ __forceinline__ bool sample_mean(const AeArgs& a, int x, int y, float& mean) {
    const FjAutoExposure& d = a.desc;
    const int px = d.meter_bounds.x1 + preview_coord(x, d.preview_width, d.meter_bounds.x2 - d.meter_bounds.x1);
    const int py = d.meter_bounds.y1 + preview_coord(y, d.preview_height, d.meter_bounds.y2 - d.meter_bounds.y1);
    if (px < d.source_bounds.x1 || px >= d.source_bounds.x2 || py < d.source_bounds.y1 || py >= d.source_bounds.y2)
        return false;
    const float* pix = reinterpret_cast<const float*>(a.src + static_cast<long long>(py - d.source_bounds.y1) * a.row_bytes) +
                       static_cast<long long>(px - d.source_bounds.x1) * a.comps;
    const float in[3] = {pix[0], pix[1], pix[2]};
    float raw[3];
    camera_raw(a.ex, uniform_table(a.ex.mallett), in, raw);
    mean = (raw[0] + raw[1] + raw[2]) * (1.0f / 3.0f);
    return isfinite(mean);
}

// Per-lane weights of one sample (metering_weights, JuicerCudaAutoExposure.cu:87-153).
__device__ __forceinline__ void meter_weights(int method, int x, int y, int w, int h, float lum, float out[kLanes]) {
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (method == FJ_METER_AVERAGE || method == FJ_METER_MEDIAN) {
        out[0] = 1.0f;
        return;
    }
    const int max_dim = w > h ? w : h;
    const float inv_max = max_dim > 0 ? 1.0f / static_cast<float>(max_dim) : 0.0f;
    if (method == FJ_METER_CENTER_WEIGHTED) {
        // build_center_weight_{x,y}_kernel (:266-312): separable Gaussian, sigma 0.2.
        const float nx = (static_cast<float>(x) / static_cast<float>(w)) - 0.5f;
        const float ny = (static_cast<float>(y) / static_cast<float>(h)) - 0.5f;
        const float ux = nx * static_cast<float>(w) * inv_max, uy = ny * static_cast<float>(h) * inv_max;
        constexpr float kSigma = 0.2f;
        out[0] = expf(-(ux * ux) / (2.0f * kSigma * kSigma)) * expf(-(uy * uy) / (2.0f * kSigma * kSigma));
        return;
    }
    const float nx = ((static_cast<float>(x) / static_cast<float>(w)) - 0.5f) * static_cast<float>(w) * inv_max;
    const float ny = ((static_cast<float>(y) / static_cast<float>(h)) - 0.5f) * static_cast<float>(h) * inv_max;
    const float radius = sqrtf(nx * nx + ny * ny);
    if (method == FJ_METER_PARTIAL) {
        out[0] = radius < 0.15f ? 1.0f : 0.0f;
        out[3] = 1.0f;
    } else if (method == FJ_METER_MATRIX) {
        const int ch = h / 5, cw = w / 5;
        if (ch <= 0 || cw <= 0) return;
        const int row = y / ch, col = x / cw;
        if (row >= 5 || col >= 5) return;
        const float dy = (static_cast<float>(row) - 2.0f) / 2.0f, dx = (static_cast<float>(col) - 2.0f) / 2.0f;
        const float dist = sqrtf(dx * dx + dy * dy) / sqrtf(2.0f);
        out[0] = 0.5f * (1.0f + cosf(3.14159265358979323846f * dist));
    } else if (method == FJ_METER_MULTI_ZONE) {
        if (radius >= 0.0f && radius < 0.05f)
            out[0] = 1.0f;
        else if (radius >= 0.05f && radius < 0.25f)
            out[1] = 1.0f;
        else if (radius >= 0.25f && radius < 0.50f)
            out[2] = 1.0f;
        out[3] = 1.0f;
    } else if (method == FJ_METER_HIGHLIGHT_WEIGHTED) {
        out[0] = lum * lum;
        out[3] = 1.0f;
    }
}

__device__ __forceinline__ float scale_from(float y) {
    // 2^(-log2(Y / 1)) with the reference's guards (:500-521).
    float ev = 0.0f;
    if (y > 0.0f && isfinite(y)) {
        const float c = -log2f(y);
        if (isfinite(c)) ev = c;
    }
    const float s = exp2f(ev);
    return isfinite(s) && s > 0.0f ? s : 1.0f;
}

// Weighted sums (or maxima for the median) of one block's samples.
__global__ __launch_bounds__(kThreads) void ae_partials_kernel(AeArgs a) {
    const int w = a.desc.preview_width, h = a.desc.preview_height;
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const bool median = a.desc.method == FJ_METER_MEDIAN;
    float sy[kLanes] = {}, sw[kLanes] = {};
    float mx = 0.0f;
    if (i < w * h) {
        const int x = i % w, y = i / w;
        float mean = 0.0f;
        if (sample_mean(a, x, y, mean)) {
            if (median) {
                mx = fmaxf(0.0f, mean);
            } else {
                float wt[kLanes];
                meter_weights(static_cast<int>(a.desc.method), x, y, w, h, mean, wt);
#pragma unroll
                for (int l = 0; l < kLanes; ++l) {
                    const float v = isfinite(wt[l]) && wt[l] > 0.0f ? wt[l] : 0.0f;
                    sy[l] = mean * v;
                    sw[l] = v;
                }
            }
        }
    }
    __shared__ float s_y[kLanes][kThreads], s_w[kLanes][kThreads];
    __shared__ float s_m[kThreads];
#pragma unroll
    for (int l = 0; l < kLanes; ++l) {
        s_y[l][threadIdx.x] = sy[l];
        s_w[l][threadIdx.x] = sw[l];
    }
    s_m[threadIdx.x] = mx;
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
#pragma unroll
            for (int l = 0; l < kLanes; ++l) {
                s_y[l][threadIdx.x] += s_y[l][threadIdx.x + stride];
                s_w[l][threadIdx.x] += s_w[l][threadIdx.x + stride];
            }
            s_m[threadIdx.x] = fmaxf(s_m[threadIdx.x], s_m[threadIdx.x + stride]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        AePartial& p = a.partials[blockIdx.x];
#pragma unroll
        for (int l = 0; l < kLanes; ++l) {
            p.sum_y[l] = s_y[l][0];
            p.sum_w[l] = s_w[l][0];
        }
        p.max_y = s_m[0];
    }
}

// One block: the partials in block order, then the scale
// (finalize_auto_exposure_from_sums_kernel, :462-522); for the median the
// maximum and a cleared histogram for the next launch.
__global__ __launch_bounds__(kThreads) void ae_finalize_kernel(AeArgs a, int blocks) {
    __shared__ float s_y[kLanes][kThreads], s_w[kLanes][kThreads];
    __shared__ float s_m[kThreads];
    float sy[kLanes] = {}, sw[kLanes] = {};
    float mx = 0.0f;
    for (int b = threadIdx.x; b < blocks; b += kThreads) {
        const AePartial& p = a.partials[b];
#pragma unroll
        for (int l = 0; l < kLanes; ++l) {
            sy[l] += p.sum_y[l];
            sw[l] += p.sum_w[l];
        }
        mx = fmaxf(mx, p.max_y);
    }
#pragma unroll
    for (int l = 0; l < kLanes; ++l) {
        s_y[l][threadIdx.x] = sy[l];
        s_w[l][threadIdx.x] = sw[l];
    }
    s_m[threadIdx.x] = mx;
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
#pragma unroll
            for (int l = 0; l < kLanes; ++l) {
                s_y[l][threadIdx.x] += s_y[l][threadIdx.x + stride];
                s_w[l][threadIdx.x] += s_w[l][threadIdx.x + stride];
            }
            s_m[threadIdx.x] = fmaxf(s_m[threadIdx.x], s_m[threadIdx.x + stride]);
        }
        __syncthreads();
    }
    const int method = static_cast<int>(a.desc.method);
    if (method == FJ_METER_MEDIAN) {
        for (int i = threadIdx.x; i < kAeBins; i += kThreads) a.histogram[i] = 0u;
        if (threadIdx.x == 0) *a.max_y = s_m[0];
        return;
    }
    if (threadIdx.x != 0) return;
    float y = 0.0f;
    bool has = false;
    auto lane_y = [&](int l) { return s_y[l][0]; };
    auto lane_w = [&](int l) { return s_w[l][0]; };
    if (method == FJ_METER_PARTIAL) {
        const int l = lane_w(0) > 0.0f ? 0 : 3;
        if (lane_w(l) > 0.0f) {
            y = lane_y(l) / lane_w(l);
            has = true;
        }
    } else if (method == FJ_METER_MULTI_ZONE) {
        const float ring[3] = {0.50f, 0.30f, 0.20f};
        float sum = 0.0f, total = 0.0f;
        for (int l = 0; l < 3; ++l) {
            if (lane_w(l) > 0.0f) {
                sum += ring[l] * lane_y(l) / lane_w(l);
                total += ring[l];
            }
        }
        if (total > 0.0f) {
            y = sum / total;
            has = true;
        } else if (lane_w(3) > 0.0f) {
            y = lane_y(3) / lane_w(3);
            has = true;
        }
    } else if (method == FJ_METER_HIGHLIGHT_WEIGHTED && lane_w(0) < 1e-12f) {
        if (lane_w(3) > 0.0f) {
            y = lane_y(3) / lane_w(3);
            has = true;
        }
    } else if (lane_w(0) > 0.0f) {
        y = lane_y(0) / lane_w(0);
        has = true;
    }
    *a.scale = has ? scale_from(y) : 1.0f;
}

// Median histogram of the raw means over [0, max] (meter_histogram_Y_bits_kernel, :524-597).
__global__ __launch_bounds__(kThreads) void ae_histogram_kernel(AeArgs a) {
    __shared__ unsigned int s_h[kAeBins];
    for (int i = threadIdx.x; i < kAeBins; i += kThreads) s_h[i] = 0u;
    __syncthreads();
    const float max_y = *a.max_y;
    const int w = a.desc.preview_width, h = a.desc.preview_height;
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (max_y > 0.0f && isfinite(max_y) && i < w * h) {
        float mean = 0.0f;
        if (sample_mean(a, i % w, i / w, mean)) {
            if (mean < 0.0f) mean = 0.0f;
            const float norm = fminf(1.0f, mean / max_y);
            const int bin = static_cast<int>(norm * static_cast<float>(kAeBins - 1));
            atomicAdd(&s_h[bin], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kAeBins; b += kThreads)
        if (s_h[b]) atomicAdd(&a.histogram[b], s_h[b]);
}

// Median of the histogram and its scale (finalize_auto_exposure_from_histogram_kernel, :599-660).
__global__ void ae_median_kernel(AeArgs a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float max_y = *a.max_y;
    if (!(max_y > 0.0f) || !isfinite(max_y)) {
        *a.scale = 1.0f;
        return;
    }
    unsigned long long total = 0;
    for (int i = 0; i < kAeBins; ++i) total += a.histogram[i];
    if (total == 0) {
        *a.scale = 1.0f;
        return;
    }
    const unsigned long long target = total / 2ull;
    unsigned long long cumulative = 0;
    int bin = kAeBins - 1;
    unsigned int in_bin = 0;
    for (int i = 0; i < kAeBins; ++i) {
        const unsigned int c = a.histogram[i];
        if (cumulative + c > target) {
            bin = i;
            in_bin = c;
            break;
        }
        cumulative += c;
    }
    const float width = max_y / static_cast<float>(kAeBins);
    const float fraction = in_bin > 0 ? static_cast<float>(target - cumulative) / static_cast<float>(in_bin) : 0.0f;
    *a.scale = scale_from((static_cast<float>(bin) + fraction) * width);
}

}  // namespace

int ae_blocks(const FjAutoExposure& d) {
    const long long n = static_cast<long long>(d.preview_width) * d.preview_height;
    return n > 0 ? static_cast<int>((n + kThreads - 1) / kThreads) : 0;
}

hipError_t launch_auto_exposure(const AeArgs& a, hipStream_t stream) {
    const int blocks = ae_blocks(a.desc);
    if (blocks <= 0 || blocks > kAeMaxBlocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ae_partials_kernel, dim3(blocks), dim3(kThreads), 0, stream, a);
    hipLaunchKernelGGL(ae_finalize_kernel, dim3(1), dim3(kThreads), 0, stream, a, blocks);
    if (a.desc.method == FJ_METER_MEDIAN) {
        hipLaunchKernelGGL(ae_histogram_kernel, dim3(blocks), dim3(kThreads), 0, stream, a);
        hipLaunchKernelGGL(ae_median_kernel, dim3(1), dim3(64), 0, stream, a);
    }
    return hipGetLastError();
}

}  // namespace fjx
