// fj_post.hip — scanner post stage: unsharp mask and output encoding fused
// into one LDS-tiled pass over the three linear RGB planes.
//
// Reference behaviour (JuicerCudaScanPipeline.cu:316-452, :586-660,
// :2096-2182): horizontal Gaussian blur, vertical blur, v + a*(v - blur),
// non-finite intermediate values replaced by zero, half-sample reflected
// borders, then the output CCTF, a [0,1] clamp and the source alpha. With a
// gate weave the unsharp mask has run before (the weave resamples its
// result) and weave_encode_kernel samples the planes at the woven positions
// in place of this pass.
#include <hip/hip_runtime.h>

#include "fj_batch.h"
#include "fj_device.h"
#include "fj_pixel_math.h"

namespace fjx {

namespace {

constexpr int kTileW = 64;
constexpr int kTileH = 16;
constexpr int kMaxPostRadius = 16;

__device__ __forceinline__ int reflect_half(int i, int n) {
    if (n <= 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * n - i - 1;
    return i;
}

// Mitchell-Netravali cubic with B = C = 1/3 (Mitchell & Netravali, 1988):
// 7|t|^3 - 12t^2 + 16/3 inside |t| < 1, -7/3|t|^3 + 12t^2 - 20|t| + 32/3 in
// 1 <= |t| < 2, both over 6.
__device__ __forceinline__ float mitchell(float t) {
    const float a = fabsf(t);
    if (a < 1.0f) return ((7.0f * a - 12.0f) * a * a + 16.0f / 3.0f) * (1.0f / 6.0f);
    if (a < 2.0f) return (((-7.0f / 3.0f) * a + 12.0f) * a - 20.0f) * a * (1.0f / 6.0f) + 32.0f / 18.0f;
    return 0.0f;
}

// One plane at the woven source position of output pixel (x, y): rotation
// about the window centre, then the shift; the position is clamped to the
// plane and the 4 x 4 taps clamp their indices (edge replication); the 16
// products are summed in FP64 (JuicerCudaScanPipeline.cu:1514-1558,
// :2122-2146). A non-finite position clamps to 0 (fmaxf drops the NaN).
__device__ __forceinline__ float weave_sample(const PostParams& p, const float* plane, int x, int y) {
    const int W = p.io.width, H = p.io.height;
    const float cx = 0.5f * static_cast<float>(W - 1), cy = 0.5f * static_cast<float>(H - 1);
    const float ox = static_cast<float>(x) - cx, oy = static_cast<float>(y) - cy;
    float sx = p.weave_cos * ox - p.weave_sin * oy + cx + p.weave_dx;
    float sy = p.weave_sin * ox + p.weave_cos * oy + cy + p.weave_dy;
    sx = fminf(fmaxf(sx, 0.0f), static_cast<float>(W - 1));
    sy = fminf(fmaxf(sy, 0.0f), static_cast<float>(H - 1));
    const int ix = static_cast<int>(floorf(sx)), iy = static_cast<int>(floorf(sy));
    const float fx = sx - static_cast<float>(ix), fy = sy - static_cast<float>(iy);
    float wx[4], wy[4];
    for (int k = 0; k < 4; ++k) {
        wx[k] = mitchell(fx + 1.0f - static_cast<float>(k));
        wy[k] = mitchell(fy + 1.0f - static_cast<float>(k));
    }
    double acc = 0.0;
    for (int j = 0; j < 4; ++j) {
        const int yy = min(max(iy + j - 1, 0), H - 1);
        const float* row = plane + static_cast<long long>(yy) * W;
        for (int i = 0; i < 4; ++i) {
            const int xx = min(max(ix + i - 1, 0), W - 1);
            acc = fma(static_cast<double>(wy[j]) * static_cast<double>(wx[i]), static_cast<double>(row[xx]), acc);
        }
    }
    return static_cast<float>(acc);
}

}  // namespace

__global__ __launch_bounds__(256) void post_kernel(PostParams p) {
    __shared__ float s_in[(kTileH + 2 * kMaxPostRadius) * (kTileW + 2 * kMaxPostRadius)];
    __shared__ float s_h[(kTileH + 2 * kMaxPostRadius) * kTileW];
    __shared__ float s_w[2 * kMaxPostRadius + 1];
    const int r = p.radius;
    const int W = p.io.width, H = p.io.height;
    const long long plane = static_cast<long long>(W) * H;
    const int x0 = blockIdx.x * kTileW, y0 = blockIdx.y * kTileH;
    const int in_w = kTileW + 2 * r, in_h = kTileH + 2 * r;
    for (int i = threadIdx.x; i < 2 * r + 1; i += blockDim.x) s_w[i] = p.weights[i];
    // each thread owns 4 pixels: column tx, rows ty*4 .. ty*4+3
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    float out[3][4];
    for (int c = 0; c < 3; ++c) {
        const float* src = p.planes + c * plane;
        if (r > 0) {
            __syncthreads();
            batched<8, float>(
                threadIdx.x, in_w * in_h, blockDim.x,
                [&](int i) {
                    const int ly = i / in_w, lx = i - ly * in_w;
                    const int gy = reflect_half(y0 + ly - r, H), gx = reflect_half(x0 + lx - r, W);
                    return src[static_cast<long long>(gy) * W + gx];
                },
                [&](int i, float v) { s_in[i] = v; });
            __syncthreads();
            for (int i = threadIdx.x; i < kTileW * in_h; i += blockDim.x) {
                const int ly = i / kTileW, lx = i - ly * kTileW;
                float acc = 0.0f;
                for (int j = 0; j <= 2 * r; ++j) acc = fmaf(s_in[ly * in_w + lx + j], s_w[j], acc);
                s_h[ly * kTileW + lx] = isfinite(acc) ? acc : 0.0f;
            }
            __syncthreads();
            for (int q = 0; q < 4; ++q) {
                const int ly = ty * 4 + q;
                float acc = 0.0f;
                for (int j = 0; j <= 2 * r; ++j) acc = fmaf(s_h[(ly + j) * kTileW + tx], s_w[j], acc);
                const float blurred = isfinite(acc) ? acc : 0.0f;
                const float v0 = s_in[(ly + r) * in_w + tx + r];
                const float v = fmaf(p.amount, v0 - blurred, v0);
                out[c][q] = isfinite(v) ? v : 0.0f;
            }
        } else {
            for (int q = 0; q < 4; ++q) {
                const int gx = x0 + tx, gy = y0 + ty * 4 + q;
                out[c][q] = (gx < W && gy < H) ? src[static_cast<long long>(gy) * W + gx] : 0.0f;
            }
        }
    }
    for (int q = 0; q < 4; ++q) {
        const int gx = x0 + tx, gy = y0 + ty * 4 + q;
        if (gx >= W || gy >= H) continue;
        float rgb[3] = {out[0][q], out[1][q], out[2][q]};
        encode_output(p.co, rgb);
        const float* sp = reinterpret_cast<const float*>(p.io.src + gy * p.io.src_row_bytes) + gx * p.io.comps;
        float* dp = reinterpret_cast<float*>(p.io.dst + gy * p.io.dst_row_bytes) + gx * p.io.comps;
        if (p.io.comps == 4) {
            *reinterpret_cast<float4*>(dp) = make_float4(rgb[0], rgb[1], rgb[2], sp[3]);
        } else {
            dp[0] = rgb[0]; dp[1] = rgb[1]; dp[2] = rgb[2];
        }
    }
}

// The weave's own pass, one pixel per thread: kept apart from post_kernel,
// whose tiles would otherwise carry the FP64 taps' registers.
__global__ __launch_bounds__(256) void weave_encode_kernel(PostParams p) {
    const long long n = static_cast<long long>(p.io.width) * p.io.height;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int gy = static_cast<int>(i / p.io.width), gx = static_cast<int>(i - static_cast<long long>(gy) * p.io.width);
    float rgb[3];
    for (int c = 0; c < 3; ++c) rgb[c] = weave_sample(p, p.planes + c * n, gx, gy);
    encode_output(p.co, rgb);
    const float* sp = reinterpret_cast<const float*>(p.io.src + gy * p.io.src_row_bytes) + gx * p.io.comps;
    float* dp = reinterpret_cast<float*>(p.io.dst + gy * p.io.dst_row_bytes) + gx * p.io.comps;
    if (p.io.comps == 4) {
        *reinterpret_cast<float4*>(dp) = make_float4(rgb[0], rgb[1], rgb[2], sp[3]);
    } else {
        dp[0] = rgb[0]; dp[1] = rgb[1]; dp[2] = rgb[2];
    }
}

hipError_t launch_post(const PostParams& p, hipStream_t stream) {
    if (p.radius < 0 || p.radius > kMaxPostRadius) return hipErrorInvalidValue;
    if (p.weave) {
        if (p.radius != 0) return hipErrorInvalidValue;
        const long long n = static_cast<long long>(p.io.width) * p.io.height;
        hipLaunchKernelGGL(weave_encode_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, p);
        return hipGetLastError();
    }
    const dim3 grid((p.io.width + kTileW - 1) / kTileW, (p.io.height + kTileH - 1) / kTileH);
    hipLaunchKernelGGL(post_kernel, grid, dim3(256), 0, stream, p);
    return hipGetLastError();
}

int max_post_radius() { return kMaxPostRadius; }

}  // namespace fjx
