// fj_glare.hip — scanner glare field generation.
//
// The reference draws each pixel's glare from openrand's Philox4x32-10
// (third_party/openrand/philox.h:64-145) in optics_glare_generate_kernel
// (src/Cuda/Scan/JuicerCudaScanPipeline.cu:273-314). Philox here is written
// from Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as
// 1, 2, 3" (SC 2011): ten rounds of the 4x32 S-box with the paper's
// multipliers and Weyl key increments. The stream arrangement (key = seed
// high/low words, counter = {x, y, medium, draw}, first output word per
// draw) and the FP32 uniform (x * 2^-32 + 2^-33) are the reference's, so the
// uniforms are bit-identical to its; the transcendental steps use this
// device's library (cospi of the exact 2u argument).
#include "fj_glare.h"

namespace fjx {
namespace {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// First output word of Philox4x32-10 for two counters that differ only in
// their last word (draws 0 and 1 of one pixel), interleaved for ILP.
__device__ __forceinline__ void philox_pair(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                            uint32_t& out0, uint32_t& out1) {
    uint32_t a0 = c0, a1 = c1, a2 = c2, a3 = 0u;
    uint32_t b0 = c0, b1 = c1, b2 = c2, b3 = 1u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint32_t ahi0 = __umulhi(kPhiloxM0, a0), alo0 = kPhiloxM0 * a0;
        const uint32_t ahi1 = __umulhi(kPhiloxM1, a2), alo1 = kPhiloxM1 * a2;
        const uint32_t bhi0 = __umulhi(kPhiloxM0, b0), blo0 = kPhiloxM0 * b0;
        const uint32_t bhi1 = __umulhi(kPhiloxM1, b2), blo1 = kPhiloxM1 * b2;
        const uint32_t an0 = ahi1 ^ a1 ^ k0, an2 = ahi0 ^ a3 ^ k1;
        const uint32_t bn0 = bhi1 ^ b1 ^ k0, bn2 = bhi0 ^ b3 ^ k1;
        a0 = an0; a1 = alo1; a2 = an2; a3 = alo0;
        b0 = bn0; b1 = blo1; b2 = bn2; b3 = blo0;
    }
    out0 = a0;
    out1 = b0;
}

__device__ __forceinline__ float u01(uint32_t x) {
    // (float)x is exact to round-to-nearest; the scale by 2^-32 is exact.
    return fmaf(static_cast<float>(x), 0x1p-32f, 0x1p-33f);
}

__global__ __launch_bounds__(256) void glare_generate_kernel(GlareArgs a) {
    const long long total = static_cast<long long>(a.width) * a.height;
    const uint32_t k0 = static_cast<uint32_t>(a.seed >> 32), k1 = static_cast<uint32_t>(a.seed);
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(i / a.width);
        const int x = static_cast<int>(i - static_cast<long long>(y) * a.width);
        const uint32_t ax = static_cast<uint32_t>(static_cast<long long>(a.origin_x) + x);
        const uint32_t ay = static_cast<uint32_t>(static_cast<long long>(a.origin_y) + y);
        uint32_t w0, w1;
        philox_pair(k0, k1, ax, ay, a.medium, w0, w1);
        const float r1 = u01(w0), u2 = u01(w1);
        if (a.uniforms) {
            a.uniforms[i] = r1;
            a.uniforms[total + i] = u2;
        }
        const float u1 = fminf(fmaxf(r1, 1e-7f), 1.0f);
        const float n = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
        const float g = expf(fmaf(a.sigma, n, a.mu));
        a.out[i] = isfinite(g) ? g : 0.0f;
    }
}

}  // namespace

hipError_t launch_glare_generate(const GlareArgs& a, hipStream_t stream) {
    if (!a.out || a.width <= 0 || a.height <= 0) return hipErrorInvalidValue;
    const long long total = static_cast<long long>(a.width) * a.height;
    const long long blocks = (total + 255) / 256;
    const unsigned int grid = static_cast<unsigned int>(blocks < 8192 ? blocks : 8192);
    glare_generate_kernel<<<grid, 256, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace fjx
