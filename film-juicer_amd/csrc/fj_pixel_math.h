// fj_pixel_math.h — FP32 per-pixel stage math shared by the fused and the
// staged kernels. Each function cites the reference behaviour it keeps.
#pragma once

#include <hip/hip_runtime.h>

#include "fj_device.h"

namespace fjx {

__device__ __forceinline__ float fin0(float v) { return isfinite(v) ? v : 0.0f; }

__device__ __forceinline__ void mat3v(const float* m, const float* v, float* o) {
    o[0] = fmaf(m[0], v[0], fmaf(m[1], v[1], m[2] * v[2]));
    o[1] = fmaf(m[3], v[0], fmaf(m[4], v[1], m[5] * v[2]));
    o[2] = fmaf(m[6], v[0], fmaf(m[7], v[1], m[8] * v[2]));
}

// log2 / exp2 on the hardware transcendental unit (1 ulp on normal inputs).
__device__ __forceinline__ float hw_log2(float v) { return __builtin_amdgcn_logf(v); }
__device__ __forceinline__ float hw_exp2(float v) { return __builtin_amdgcn_exp2f(v); }
__device__ __forceinline__ float log10_fast(float v) { return hw_log2(v) * 0.30102999566398119521f; }

// Input CCTF decode (JuicerCudaDeviceHelpers.cuh:179-225).
__device__ __forceinline__ float decode_channel(int space, float v) {
    const float x = fmaxf(0.0f, fin0(v));
    if (space == 1) {  // BT.2020
        return x < 0.0812428791f ? x / 4.5f : powf((x + 0.09929681f) / 1.09929681f, 1.0f / 0.45f);
    }
    if (space == 3) {  // sRGB
        return x <= 0.04045f ? x / 12.92f : powf((x + 0.055f) / 1.055f, 2.4f);
    }
    return fin0(v);
}

// Mitchell–Netravali B = C = 1/3 weights for the four taps of fraction f
// (JuicerCudaDeviceHelpers.cuh:266-282), as polynomials in f.
__device__ __forceinline__ void mitchell4(float f, float w[4]) {
    const float f2 = f * f, f3 = f2 * f;
    // |t| in [1,2) branch for t = f+1 and t = f-2 (mirrored), [0,1) for t = f, f-1.
    const float a = 1.0f + f, b = 2.0f - f;
    const float k1 = -7.0f / 18.0f, k2 = 2.0f, k3 = -10.0f / 3.0f, k4 = 16.0f / 9.0f;
    const float m1 = 7.0f / 6.0f, m2 = -2.0f, m3 = 8.0f / 9.0f;
    w[0] = fmaf(fmaf(fmaf(k1, a, k2), a, k3), a, k4);
    w[3] = fmaf(fmaf(fmaf(k1, b, k2), b, k3), b, k4);
    w[1] = fmaf(m1, f3, fmaf(m2, f2, m3));
    const float g = 1.0f - f, g2 = g * g, g3 = g2 * g;
    w[2] = fmaf(m1, g3, fmaf(m2, g2, m3));
}

__device__ __forceinline__ int reflect192(int i) {
    return i < 0 ? -i : (i >= kTcExtent ? 2 * (kTcExtent - 1) - i : i);
}

__device__ __forceinline__ void tc_coord(float q, int& base, float& frac) {
    const float c = fminf(1.0f, fmaxf(0.0f, q)) * static_cast<float>(kTcExtent - 1);
    if (c >= static_cast<float>(kTcExtent - 1)) {
        base = kTcExtent - 2;
        frac = 1.0f;
    } else {
        base = static_cast<int>(floorf(c));
        frac = c - static_cast<float>(base);
    }
}

// A wave-uniform table (the spectral loops read entry i for every lane at
// once) viewed through the constant address space: loads at a uniform index
// become scalar loads into SGPRs, with no LDS staging and no VGPRs.
typedef __attribute__((address_space(4))) const float UniformF;
struct UniformF4 {
    const UniformF* f;  // 4 floats per entry
};
__device__ __forceinline__ UniformF4 uniform_table(const float4* p) {
    return UniformF4{(const UniformF*)p};  // address-space cast (a C-style cast is the only form allowed)
}
__device__ __forceinline__ float4 uniform_at(UniformF4 t, int i) {
    return make_float4(t.f[4 * i], t.f[4 * i + 1], t.f[4 * i + 2], t.f[4 * i + 3]);
}

// Mallett 2019 spectral reconstruction exposure in B/G/R
// (JuicerCudaDeviceHelpers.cuh:619-668, :699-703).
__device__ __forceinline__ void mallett_exposure(const ExposureParams& ex, UniformF4 table, const float rgb[3],
                                                 float E[3]) {
    float s[3];
    mat3v(ex.rgb_to_srgb, rgb, s);
    const float r = fmaxf(0.0f, fin0(s[0])), g = fmaxf(0.0f, fin0(s[1])), b = fmaxf(0.0f, fin0(s[2]));
    float eb = 0.0f, eg = 0.0f, er = 0.0f;
#pragma unroll 4
    for (int i = 0; i < ex.mallett_k; ++i) {
        const float4 w = uniform_at(table, 2 * i);
        const float4 sens = uniform_at(table, 2 * i + 1);
        const float spd = fmaf(r, w.x, fmaf(g, w.y, b * w.z));
        eb = fmaf(spd, sens.x, eb);
        eg = fmaf(spd, sens.y, eg);
        er = fmaf(spd, sens.z, er);
    }
    E[0] = fmaxf(0.0f, (isfinite(eb) ? fmaxf(0.0f, eb) : 0.0f) * ex.mallett_scale);
    E[1] = fmaxf(0.0f, (isfinite(eg) ? fmaxf(0.0f, eg) : 0.0f) * ex.mallett_scale);
    E[2] = fmaxf(0.0f, (isfinite(er) ? fmaxf(0.0f, er) : 0.0f) * ex.mallett_scale);
}

// Reconstructed film raw exposure in B/G/R before any exposure scale
// (reconstruct_film_raw_device, JuicerCudaDeviceHelpers.cuh:509-617,
// :670-708): XYZ -> chromaticity -> tri2quad -> bicubic TC LUT -> x sum;
// Mallett inputs take the spectral path above.
__device__ __forceinline__ void camera_raw(const ExposureParams& ex, UniformF4 mallett_table, const float in[3],
                                           float E[3]) {
    float rgb[3];
    if (ex.decode_cctf) {
        rgb[0] = decode_channel(ex.decode_cctf, in[0]);
        rgb[1] = decode_channel(ex.decode_cctf, in[1]);
        rgb[2] = decode_channel(ex.decode_cctf, in[2]);
    } else {
        rgb[0] = fin0(in[0]); rgb[1] = fin0(in[1]); rgb[2] = fin0(in[2]);
    }
    if (ex.mallett) {
        mallett_exposure(ex, mallett_table, rgb, E);
        return;
    }
    float w[3];
    mat3v(ex.rgb_to_xyz, rgb, w);
    w[0] = fin0(w[0]); w[1] = fin0(w[1]); w[2] = fin0(w[2]);
    const float sum = w[0] + w[1] + w[2];
    const float inv = 1.0f / fmaxf(sum, 1e-10f);
    const float cx = w[0] * inv, cy = w[1] * inv;
    const float omx = 1.0f - cx;
    const float qx = fminf(1.0f, fmaxf(0.0f, omx * omx));
    const float qy = fminf(1.0f, fmaxf(0.0f, cy / fmaxf(omx, 1e-10f)));
    int bc, bm;
    float fc, fm;
    tc_coord(qx, bc, fc);
    tc_coord(qy, bm, fm);
    float wc[4], wm[4];
    mitchell4(fc, wc);
    mitchell4(fm, wm);
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
    float wsum = 0.0f;
#pragma unroll
    for (int dc = 0; dc < 4; ++dc) {
        const int c = reflect192(bc - 1 + dc);
        const float4* row = ex.tc_lut + c * kTcExtent;
        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, rs = 0.0f;
#pragma unroll
        for (int dm = 0; dm < 4; ++dm) {
            const int m = reflect192(bm - 1 + dm);
            const float4 t = __ldg(row + m);
            r0 = fmaf(wm[dm], t.x, r0);
            r1 = fmaf(wm[dm], t.y, r1);
            r2 = fmaf(wm[dm], t.z, r2);
            rs += wm[dm];
        }
        acc0 = fmaf(wc[dc], r0, acc0);
        acc1 = fmaf(wc[dc], r1, acc1);
        acc2 = fmaf(wc[dc], r2, acc2);
        wsum = fmaf(wc[dc], rs, wsum);
    }
    const float winv = wsum != 0.0f ? 1.0f / wsum : 0.0f;
    const float lr = fin0(acc0 * winv), lg = fin0(acc1 * winv), lb = fin0(acc2 * winv);
    E[0] = fin0(sum * lb);
    E[1] = fin0(sum * lg);
    E[2] = fin0(sum * lr);
}

// The exposure scale of a launch (compute_film_linear_exposure_device,
// JuicerCudaDeviceHelpers.cuh:709-744): with auto exposure, the metered scale
// times the manual scale (times the route correction when the launch applies
// it), else the launch's precomputed scale.
__device__ __forceinline__ float exposure_scale(const ExposureParams& ex) {
    if (!ex.ae) return ex.scale;
    float a = *ex.ae;
    if (!(isfinite(a) && a > 0.0f)) a = 1.0f;
    float s = a * ex.camera_scale;
    if (ex.with_route) s *= ex.route_scale;
    return isfinite(s) && s > 0.0f ? s : 1.0f;
}

// Camera film linear exposure in B/G/R: the raw exposure times the scale.
__device__ __forceinline__ void camera_exposure(const ExposureParams& ex, UniformF4 mallett_table,
                                                const float in[3], float E[3]) {
    camera_raw(ex, mallett_table, in, E);
    const float s = exposure_scale(ex);
    E[0] = fmaxf(0.0f, E[0] * s);
    E[1] = fmaxf(0.0f, E[1] * s);
    E[2] = fmaxf(0.0f, E[2] * s);
}

// Clamped linear lookup on an ascending axis (JuicerCudaDeviceHelpers.cuh:34-89).
// xq must be finite. Branch-free bisection for the last x[i] <= xq.
__device__ __forceinline__ float curve_at(const float* x, const float* y, int n, float xq) {
    const float xmin = x[0], xmax = x[n - 1];
    if (!(xmax >= xmin) || xq <= xmin) return y[0];
    if (xq >= xmax) return y[n - 1];
    int i = 0;
    int step = 1;
    while (step * 2 < n) step *= 2;
    for (; step > 0; step >>= 1) {
        const int j = i + step;
        if (j < n && x[j] <= xq) i = j;
    }
    if (i >= n - 1) return y[n - 1];
    const float x0 = x[i], x1 = x[i + 1], y0 = y[i], y1 = y[i + 1];
    const float den = x1 - x0;
    if (!(den > 0.0f) || !isfinite(den)) return y0;
    return fmaf((xq - x0) / den, y1 - y0, y0);
}

__device__ __forceinline__ float sample_logE(const float* x, const float* y, int n, float logE, float gamma) {
    if (!isfinite(logE)) return __builtin_nanf("");
    return curve_at(x, y, n, logE * gamma);
}

// Guarded Langmuir v*(k+ref)/(k+v) (JuicerCudaDeviceHelpers.cuh:1073-1123).
__device__ __forceinline__ bool langmuir(float v, float k, float ref, float& out) {
    out = __builtin_nanf("");
    if (!isfinite(v) || !isfinite(k) || !(k > 0.0f) || !isfinite(ref)) return false;
    const float den = k + v;
    const float kr = k + ref;
    if (isfinite(den) && den != 0.0f && isfinite(kr) && (v == 0.0f || fabsf(v) <= 3.402823466e38f / fabsf(kr))) {
        out = v * kr / den;
        return isfinite(out);
    }
    const float ns = 1.0f + ref / k;
    if (!isfinite(ns)) return false;
    if (v >= 0.0f) {
        const float ds = 1.0f + v / k;
        if (!isfinite(ds) || ds == 0.0f) return false;
        out = (v / ds) * ns;
        return isfinite(out);
    }
    if (!isfinite(den) || den == 0.0f) return false;
    const float ds = den / k;
    if (!isfinite(ds) || ds == 0.0f) return false;
    out = fabsf(v) <= 3.402823466e38f / fabsf(ns) ? v * ns / ds : (v / ds) * ns;
    return isfinite(out);
}

// Film development with point DIR couplers; D out in C/M/Y. Returns DIR
// failure bits (JuicerCudaScanPipeline.cu:1624-1697,
// JuicerCudaDeviceHelpers.cuh:1125-1253).
__device__ __forceinline__ unsigned int develop_film(const DevelopParams& dv, const float* curves, const float E[3],
                                                     float D[3]) {
    const int n = dv.n;
    const float* x = curves;
    float logE[3], layer[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        float l = log10_fast(fmaxf(E[b], 0.0f) + 1e-10f);
        if (isinf(l)) l = (x[n - 1] >= x[0]) ? (l > 0.0f ? x[n - 1] : x[0]) : l;
        logE[b] = l;
        layer[b] = sample_logE(x, curves + (1 + b) * n, n, l, dv.gamma_bgr[b]);
    }
    if (dv.dir_mode == 0) {
        D[0] = layer[2];
        D[1] = layer[1];
        D[2] = layer[0];
        return 0u;
    }
    unsigned int fail = 0u;
    float donors[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        donors[c] = dv.dmax[c] - layer[c];
        if (dv.dir_mode == 1 && !langmuir(layer[c], dv.donor_k[c], dv.dref[c], donors[c])) fail |= 1u << c;
    }
    float corr[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        corr[r] = fmaf(dv.M[r], donors[0], fmaf(dv.M[3 + r], donors[1], dv.M[6 + r] * donors[2]));
        if (!isfinite(corr[r])) fail |= 1u << r;
    }
    if (dv.dir_mode == 2) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (dv.recv_cref[c] == 0.0f) {
                corr[c] = 0.0f;
                continue;
            }
            float v;
            if (!langmuir(corr[c], dv.recv_kr[c], dv.recv_cref[c], v)) fail |= 1u << c;
            corr[c] = v;
        }
    }
    if (fail == 0u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float o = logE[c];
            if (isnan(o)) {
                fail |= 1u << c;
                continue;
            }
            logE[c] = o - corr[c];
            if ((!isinf(o) && !isfinite(logE[c])) ||
                (isinf(o) && (!isinf(logE[c]) || signbit(o) != signbit(logE[c]))))
                fail |= 1u << c;
        }
    }
    if (fail) {
        D[0] = D[1] = D[2] = __builtin_nanf("");
        return (1u << 8) | ((fail & 7u) << 9);
    }
    D[2] = sample_logE(curves + 4 * n, curves + 1 * n, n, logE[0], dv.gamma_bgr[0]);
    D[1] = sample_logE(curves + 5 * n, curves + 2 * n, n, logE[1], dv.gamma_bgr[1]);
    D[0] = sample_logE(curves + 6 * n, curves + 3 * n, n, logE[2], dv.gamma_bgr[2]);
    return 0u;
}

// Spatial-DIR source: raw donor corrections (Y/M/C) and the unsanitized log
// raw exposure (B/G/R) of one pixel (JuicerCudaFilmPipeline.cu:124-220).
__device__ __forceinline__ unsigned int dir_source(const DevelopParams& dv, const float* curves, const float E[3],
                                                   float corr[3], float log_raw[3]) {
    const int n = dv.n;
    const float* x = curves;
    float layer[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const float raw = log10_fast(fmaxf(E[b], 0.0f) + 1e-10f);
        log_raw[b] = raw;
        float l = raw;
        if (isinf(l)) l = (x[n - 1] >= x[0]) ? (l > 0.0f ? x[n - 1] : x[0]) : l;
        layer[b] = sample_logE(x, curves + (1 + b) * n, n, l, dv.gamma_bgr[b]);
    }
    unsigned int fail = 0u;
    float donors[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        donors[c] = dv.dmax[c] - layer[c];
        if (dv.dir_mode == 1 && !langmuir(layer[c], dv.donor_k[c], dv.dref[c], donors[c])) fail |= 1u << c;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        corr[r] = fmaf(dv.M[r], donors[0], fmaf(dv.M[3 + r], donors[1], dv.M[6 + r] * donors[2]));
        if (!isfinite(corr[r])) fail |= 1u << r;
    }
    return fail ? ((1u << 8) | ((fail & 7u) << 9)) : 0u;
}

// Final development against filtered corrections
// (JuicerCudaDeviceHelpers.cuh:1255-1329). D out in C/M/Y.
__device__ __forceinline__ unsigned int develop_dir_final(const DevelopParams& dv, const float* curves,
                                                          const float log_raw[3], const float filtered[3], float D[3]) {
    const int n = dv.n;
    const float* x = curves;
    float corr[3] = {filtered[0], filtered[1], filtered[2]};
    unsigned int fail = 0u;
    if (dv.dir_mode == 2) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (dv.recv_cref[c] == 0.0f) {
                corr[c] = 0.0f;
                continue;
            }
            float v;
            if (!langmuir(corr[c], dv.recv_kr[c], dv.recv_cref[c], v)) fail |= 1u << c;
            corr[c] = v;
        }
    }
    float l[3];
    if (fail == 0u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float o = log_raw[c];
            if (isnan(o)) {
                fail |= 1u << c;
                continue;
            }
            l[c] = o - corr[c];
            if ((!isinf(o) && !isfinite(l[c])) || (isinf(o) && (!isinf(l[c]) || signbit(o) != signbit(l[c]))))
                fail |= 1u << c;
        }
    }
    if (fail) {
        D[0] = D[1] = D[2] = __builtin_nanf("");
        return (1u << 8) | ((fail & 7u) << 9);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (isinf(l[c])) l[c] = (x[n - 1] >= x[0]) ? (l[c] > 0.0f ? x[n - 1] : x[0]) : l[c];
    D[2] = sample_logE(curves + 4 * n, curves + 1 * n, n, l[0], dv.gamma_bgr[0]);
    D[1] = sample_logE(curves + 5 * n, curves + 2 * n, n, l[1], dv.gamma_bgr[1]);
    D[0] = sample_logE(curves + 6 * n, curves + 3 * n, n, l[2], dv.gamma_bgr[2]);
    return 0u;
}

// 81-λ print exposure over the compacted active wavelengths, then the print
// curves (JuicerCudaDeviceHelpers.cuh:866-1071). A NaN transmitted sample
// contributes zero, as in the reference.
__device__ __forceinline__ void print_inner(const PrintParams& pr, UniformF4 spec, const float D[3],
                                            float inner[3]) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll 4
    for (int i = 0; i < pr.k; ++i) {
        const float4 e = uniform_at(spec, 2 * i);
        const float4 s = uniform_at(spec, 2 * i + 1);
        const float ex = fmaf(D[0], e.x, fmaf(D[1], e.y, fmaf(D[2], e.z, e.w)));
        float t = hw_exp2(ex);
        t = (t == t) ? t : 0.0f;
        a0 = fmaf(t, s.x, a0);
        a1 = fmaf(t, s.y, a1);
        a2 = fmaf(t, s.z, a2);
    }
    inner[0] = fmaf(a0, pr.normalizer, pr.preflash[0]);
    inner[1] = fmaf(a1, pr.normalizer, pr.preflash[1]);
    inner[2] = fmaf(a2, pr.normalizer, pr.preflash[2]);
}

__device__ __forceinline__ void print_develop(const PrintParams& pr, UniformF4 spec, const float* curves,
                                              float D[3]) {
    float inner[3];
    print_inner(pr, spec, D, inner);
    const int n = pr.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float l = log10_fast(inner[c] * pr.outer_scale + 1e-10f);
        D[c] = sample_logE(curves, curves + (1 + c) * n, n, l, 1.0f);
    }
}

// Enlarger-linear print exposure, the quantity the enlarger diffusion blurs
// (print_linear_for_diffusion_device, JuicerCudaDeviceHelpers.cuh:997-1010).
__device__ __forceinline__ void print_linear(const PrintParams& pr, UniformF4 spec, const float D[3],
                                             float lin[3]) {
    float inner[3];
    print_inner(pr, spec, D, inner);
#pragma unroll
    for (int c = 0; c < 3; ++c) lin[c] = (fmaxf(inner[c], 0.0f) + 1e-10f) * pr.outer_scale;
}

// Print development from the diffused print-linear exposure
// (develop_print_density_from_enlarger_linear_kernel,
// JuicerCudaScanPipeline.cu:1457-1499).
__device__ __forceinline__ void print_develop_linear(const PrintParams& pr, const float* curves, const float lin[3],
                                                     float D[3]) {
    const int n = pr.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float l = log10_fast(fmaxf(lin[c], 0.0f) + 1e-10f);
        D[c] = sample_logE(curves, curves + (1 + c) * n, n, l, 1.0f);
    }
}

// Tricubic PCHIP sample of the interleaved scanner LUT
// (JuicerCudaDeviceHelpers.cuh:359-506).
__device__ __forceinline__ float hermite(float y0, float y1, float m0, float m1, float t) {
    const float t2 = t * t, t3 = t2 * t;
    const float h00 = fmaf(2.0f, t3, fmaf(-3.0f, t2, 1.0f));
    const float h10 = fmaf(-2.0f, t2, t3 + t);
    const float h01 = fmaf(-2.0f, t3, 3.0f * t2);
    const float h11 = t3 - t2;
    return fmaf(h00, y0, fmaf(h10, m0, fmaf(h01, y1, h11 * m1)));
}

__device__ __forceinline__ void pchip_coord(float v, int res, int& b, float& t) {
    float c = v * static_cast<float>(res - 1);
    c = fminf(static_cast<float>(res - 1), fmaxf(0.0f, c));
    if (c >= static_cast<float>(res - 1)) {
        b = res - 2;
        t = 1.0f;
    } else {
        b = static_cast<int>(floorf(c));
        t = c - static_cast<float>(b);
    }
}

struct LutNode {
    float v[3], sc[3], sm[3], sy[3];
};

__device__ __forceinline__ LutNode load_node(const float4* nodes, int idx) {
    const float4 a = __ldg(nodes + 3 * idx), b = __ldg(nodes + 3 * idx + 1), c = __ldg(nodes + 3 * idx + 2);
    LutNode n;
    n.v[0] = a.x; n.v[1] = a.y; n.v[2] = a.z; n.sc[0] = a.w;
    n.sc[1] = b.x; n.sc[2] = b.y; n.sm[0] = b.z; n.sm[1] = b.w;
    n.sm[2] = c.x; n.sy[0] = c.y; n.sy[1] = c.z; n.sy[2] = c.w;
    return n;
}

__device__ __forceinline__ bool scan_lut(const ScanParams& sc, const float nrm[3], float out[3]) {
    if (!isfinite(nrm[0]) || !isfinite(nrm[1]) || !isfinite(nrm[2])) return false;
    const int r = sc.res;
    int c, m, y;
    float tc, tm, ty;
    pchip_coord(nrm[0], r, c, tc);
    pchip_coord(nrm[1], r, m, tm);
    pchip_coord(nrm[2], r, y, ty);
    const int base = (c * r + m) * r + y;
    // corners: index offsets for (dc, dm, dy)
    const LutNode n000 = load_node(sc.nodes, base);
    const LutNode n100 = load_node(sc.nodes, base + r * r);
    const LutNode n010 = load_node(sc.nodes, base + r);
    const LutNode n110 = load_node(sc.nodes, base + r * r + r);
    const LutNode n001 = load_node(sc.nodes, base + 1);
    const LutNode n101 = load_node(sc.nodes, base + r * r + 1);
    const LutNode n011 = load_node(sc.nodes, base + r + 1);
    const LutNode n111 = load_node(sc.nodes, base + r * r + r + 1);
    const int s = r - 1;
    const int cell = (c * s + m) * s + y;
    const float4 cl0 = __ldg(sc.cells + 2 * cell), cl1 = __ldg(sc.cells + 2 * cell + 1);
    const float lo[3] = {cl0.x, cl0.y, cl0.z};
    const float hi[3] = {cl0.w, cl1.x, cl1.y};
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        const float v00 = hermite(n000.v[o], n100.v[o], n000.sc[o], n100.sc[o], tc);
        const float v10 = hermite(n010.v[o], n110.v[o], n010.sc[o], n110.sc[o], tc);
        const float v01 = hermite(n001.v[o], n101.v[o], n001.sc[o], n101.sc[o], tc);
        const float v11 = hermite(n011.v[o], n111.v[o], n011.sc[o], n111.sc[o], tc);
        const float sm00 = fmaf(tc, n100.sm[o] - n000.sm[o], n000.sm[o]);
        const float sm10 = fmaf(tc, n110.sm[o] - n010.sm[o], n010.sm[o]);
        const float sm01 = fmaf(tc, n101.sm[o] - n001.sm[o], n001.sm[o]);
        const float sm11 = fmaf(tc, n111.sm[o] - n011.sm[o], n011.sm[o]);
        const float vy0 = hermite(v00, v10, sm00, sm10, tm);
        const float vy1 = hermite(v01, v11, sm01, sm11, tm);
        const float a0 = fmaf(tc, n100.sy[o] - n000.sy[o], n000.sy[o]);
        const float a1 = fmaf(tc, n110.sy[o] - n010.sy[o], n010.sy[o]);
        const float b0 = fmaf(tc, n101.sy[o] - n001.sy[o], n001.sy[o]);
        const float b1 = fmaf(tc, n111.sy[o] - n011.sy[o], n011.sy[o]);
        const float sy0 = fmaf(tm, a1 - a0, a0);
        const float sy1 = fmaf(tm, b1 - b0, b0);
        const float val = hermite(vy0, vy1, sy0, sy1, ty);
        out[o] = fminf(hi[o], fmaxf(lo[o], val));
    }
    return true;
}

// OkLch output gamut compression (src/GamutCompression.h:117-213).
__device__ __forceinline__ float knee(float v, const float k[3]) {
    if (!(v > k[0])) return v;
    const float span = k[1] - k[0];
    const float nrm = (v - k[0]) / span;
    return k[0] + span * nrm / powf(1.0f + powf(nrm, k[2]), 1.0f / k[2]);
}

__device__ __forceinline__ float gamut_boundary(const float* cmax, float L, float hue) {
    const float kPi = 3.14159265358979323846f, kTwoPi = 2.0f * kPi;
    const float cl = fminf(fmaxf(L, 0.02f), 1.0f);
    const float lp = (cl - 0.02f) / (1.0f - 0.02f) * static_cast<float>(kGamutL - 1);
    const int l0 = static_cast<int>(floorf(lp));
    const int l1 = l0 + 1 < kGamutL ? l0 + 1 : l0;
    const float lf = lp - static_cast<float>(l0);
    float wh = fmodf(hue + kPi, kTwoPi);
    if (wh < 0.0f) wh += kTwoPi;
    const float hp = wh / kTwoPi * static_cast<float>(kGamutH);
    const float hfl = floorf(hp);
    const int h0 = static_cast<int>(hfl) % kGamutH;
    const int h1 = (h0 + 1) % kGamutH;
    const float hf = hp - hfl;
    const float a0 = __ldg(cmax + l0 * kGamutH + h0), a1 = __ldg(cmax + l0 * kGamutH + h1);
    const float b0 = __ldg(cmax + l1 * kGamutH + h0), b1 = __ldg(cmax + l1 * kGamutH + h1);
    const float first = fmaf(hf, a1 - a0, a0);
    const float second = fmaf(hf, b1 - b0, b0);
    return fmaf(lf, second - first, first);
}

__device__ __forceinline__ float scbrt(float v) { return v < 0.0f ? -cbrtf(-v) : cbrtf(v); }

__device__ __forceinline__ void gamut_compress(const ColorParams& co, float rgb[3]) {
    const float kXyzLms[9] = {0.8189330101f, 0.3618667424f, -0.1288597137f, 0.0329845436f, 0.9293118715f,
                              0.0361456387f, 0.0482003018f, 0.2643662691f, 0.6338517070f};
    const float kLmsXyz[9] = {1.2270138511f, -0.5577999807f, 0.2812561490f, -0.0405801784f, 1.1122568696f,
                              -0.0716766787f, -0.0763812845f, -0.4214819784f, 1.5861632204f};
    const float kLmsLab[9] = {0.2104542553f, 0.7936177850f, -0.0040720468f, 1.9779984951f, -2.4285922050f,
                              0.4505937099f, 0.0259040371f, 0.7827717662f, -0.8086757660f};
    const float kLabLms[9] = {1.0f, 0.3963377774f, 0.2158037573f, 1.0f, -0.1055613458f,
                              -0.0638541728f, 1.0f, -0.0894841775f, -1.2914855480f};
    float xyz[3], lms[3], lab[3];
    mat3v(co.native_to_xyz, rgb, xyz);
    mat3v(kXyzLms, xyz, lms);
    lms[0] = scbrt(lms[0]); lms[1] = scbrt(lms[1]); lms[2] = scbrt(lms[2]);
    mat3v(kLmsLab, lms, lab);
    lab[0] = knee(lab[0], co.lknee);
    const float chroma = hypotf(lab[1], lab[2]);
    if (chroma > 0.0f) {
        const float hue = atan2f(lab[2], lab[1]);
        const float bnd = fmaxf(gamut_boundary(co.cmax, lab[0], hue), 1.0e-9f);
        const float comp = knee(chroma / bnd, co.cknee) * bnd;
        const float s = comp / chroma;
        lab[1] *= s;
        lab[2] *= s;
    }
    mat3v(kLabLms, lab, lms);
    lms[0] = lms[0] * lms[0] * lms[0];
    lms[1] = lms[1] * lms[1] * lms[1];
    lms[2] = lms[2] * lms[2] * lms[2];
    mat3v(kLmsXyz, lms, xyz);
    mat3v(co.xyz_to_native, xyz, rgb);
}

// Scanner: normalize -> LUT -> exp2 -> correction -> glare -> colour ->
// gamut (JuicerCudaScanPipeline.cu:1735-1811). `glare` is the pixel's glare
// percent (0 without glare). False on a scan error.
__device__ __forceinline__ bool scan_colour(const ScanParams& sc, const ColorParams& co, const float D[3],
                                            float glare, float rgb[3]) {
    float nrm[3], l2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nrm[c] = (D[c] - sc.dmin[c]) * sc.inv_span[c];
    if (!scan_lut(sc, nrm, l2)) return false;
    float xyz[3] = {hw_exp2(l2[0]), hw_exp2(l2[1]), hw_exp2(l2[2])};
    if (sc.correction) {
        const float cy = fminf(fmaxf(fmaf(sc.corr_slope, xyz[1], sc.corr_offset), 0.0f), 1.0f);
        const float s = cy / (xyz[1] + 1e-10f);
        xyz[0] *= s; xyz[1] *= s; xyz[2] *= s;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) xyz[c] = fmaf(glare, sc.glare_xyz[c], xyz[c]);
    mat3v(sc.xyz_to_rgb, xyz, rgb);
    if (!isfinite(rgb[0]) || !isfinite(rgb[1]) || !isfinite(rgb[2])) return false;
    if (co.gamut) {
        gamut_compress(co, rgb);
        if (!isfinite(rgb[0]) || !isfinite(rgb[1]) || !isfinite(rgb[2])) return false;
    }
    return true;
}

// Output CCTFs (JuicerCudaScanPipeline.cu:99-150).
__device__ __forceinline__ float pow_pos(float v, float e) { return hw_exp2(hw_log2(v) * e); }

__device__ __forceinline__ float encode_channel(const ColorParams& co, float v) {
    switch (co.cctf) {
        case 0: return v;
        case 1: return copysignf(powf(fabsf(v), co.gamma), v);
        case 2: return v <= 0.0031308f ? 12.92f * v : fmaf(1.055f, pow_pos(v, 1.0f / 2.4f), -0.055f);
        case 3: return v < co.b ? v * 4.5f : co.a * powf(v, 0.45f) - (co.a - 1.0f);
        case 4: return v < co.cutoff ? v * 16.0f : powf(v, co.gamma);
        case 5: {
            const float l = fmaxf(0.0f, v);
            return l <= co.cutoff ? l * co.d : (log2f(l + co.a) + co.b) * co.c;
        }
        default: return fminf(1.0f, fmaxf(0.0f, v));
    }
}

__device__ __forceinline__ float clamp01_keep_nan(float v) { return v <= 0.0f ? 0.0f : (v >= 1.0f ? 1.0f : v); }

__device__ __forceinline__ void encode_output(const ColorParams& co, float rgb[3]) {
    float lin[3];
    if (co.matrix) {
        mat3v(co.to_output, rgb, lin);
    } else {
        lin[0] = rgb[0]; lin[1] = rgb[1]; lin[2] = rgb[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float v = co.encode ? encode_channel(co, lin[c]) : lin[c];
        rgb[c] = clamp01_keep_nan(v);
    }
}

}  // namespace fjx
