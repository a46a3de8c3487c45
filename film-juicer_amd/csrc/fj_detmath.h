// fj_detmath.h — deterministic FP32 exp/log/sin/cos for the grain stage.
//
// Cody–Waite argument reduction (two-part ln 2 and pi/2) and Taylor
// polynomials in Horner form with explicit fused multiply-adds: every step is
// an IEEE-754 operation (fma, mul, add, division, ldexp, frexp, rint), so host
// and device evaluate the same bits. Accuracy: exp/log within 2 ulp of the
// correctly rounded value, sin/cos within 2e-7 absolute on [0, 2 pi].
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

namespace fjx {

__host__ __device__ inline float det_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    const float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-k, 0.693145751953125f, x);
    r = fmaf(-k, 1.428606765330187045037746429443359375e-06f, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, static_cast<int>(k));
}

__host__ __device__ inline float det_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    float m = frexpf(x, &e);
    if (m < 0.707106769084930419921875f) {
        m *= 2.0f;
        e -= 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = fmaf(z, 2.0f / 9.0f, 2.0f / 7.0f);
    p = fmaf(p, z, 2.0f / 5.0f);
    p = fmaf(p, z, 2.0f / 3.0f);
    const float lm = fmaf(p * z, s, 2.0f * s);
    const float ef = static_cast<float>(e);
    return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.428606765330187045037746429443359375e-06f, lm));
}

__host__ __device__ inline void det_sincosf(float x, float* sin_out, float* cos_out) {
    const float k = rintf(x * 0.636619746685028076171875f);
    float r = fmaf(-k, 1.57079637050628662109375f, x);
    r = fmaf(-k, -4.37113900018624283e-08f, r);
    const float z = r * r;
    float sp = fmaf(z, 1.0f / 362880.0f, -1.0f / 5040.0f);
    sp = fmaf(sp, z, 1.0f / 120.0f);
    sp = fmaf(sp, z, -1.0f / 6.0f);
    const float s = fmaf(sp * z, r, r);
    float cp = fmaf(z, 1.0f / 40320.0f, -1.0f / 720.0f);
    cp = fmaf(cp, z, 1.0f / 24.0f);
    cp = fmaf(cp, z, -0.5f);
    const float c = fmaf(cp, z, 1.0f);
    switch (static_cast<int>(k) & 3) {
        case 0: *sin_out = s; *cos_out = c; break;
        case 1: *sin_out = c; *cos_out = -s; break;
        case 2: *sin_out = -s; *cos_out = -c; break;
        default: *sin_out = -c; *cos_out = s; break;
    }
}

__host__ __device__ inline float det_cosf(float x) {
    float s, c;
    det_sincosf(x, &s, &c);
    return c;
}

__host__ __device__ inline float det_sinf(float x) {
    float s, c;
    det_sincosf(x, &s, &c);
    return s;
}

__host__ __device__ inline uint64_t splitmix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

}  // namespace fjx
