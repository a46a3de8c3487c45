// fj_post.hip — scanner post stage: unsharp mask and output encoding fused
// into one LDS-tiled pass over the three linear RGB planes.
//
// Reference behaviour (JuicerCudaScanPipeline.cu:316-452, :586-660,
// :2096-2182): horizontal Gaussian blur, vertical blur, v + a*(v - blur),
// non-finite intermediate values replaced by zero, half-sample reflected
// borders, then the output CCTF, a [0,1] clamp and the source alpha.
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

hipError_t launch_post(const PostParams& p, hipStream_t stream) {
    if (p.radius < 0 || p.radius > kMaxPostRadius) return hipErrorInvalidValue;
    const dim3 grid((p.io.width + kTileW - 1) / kTileW, (p.io.height + kTileH - 1) / kTileH);
    hipLaunchKernelGGL(post_kernel, grid, dim3(256), 0, stream, p);
    return hipGetLastError();
}

int max_post_radius() { return kMaxPostRadius; }

}  // namespace fjx
