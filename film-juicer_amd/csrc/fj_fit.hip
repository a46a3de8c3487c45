// fj_fit.hip — forward tail and adjoint kernels of the differentiable print
// route for gfx950.
//
// The forward reuses the render's DIR source, DCT-domain filter and DIR final
// stages, then adds the fixed grain increment and runs the print + scan on
// its own (the print curves read at print_gamma * log E). The backward runs
// the chain in reverse with three adjoint kernels and the DIR operator:
//   print adjoint   recomputes the 81-λ print integral from the saved
//                   densities (the spectra are never stored) and pushes the
//                   output gradient back through the scanner LUT, colour and
//                   OkLch gamut with forward-mode tangents (three per value:
//                   the Jacobian of the tail with respect to the scanner's
//                   normalized densities), then through the print curves, the
//                   log and the spectral sum to the densities, the print
//                   exposure, the Y/M filtration and the print gamma;
//   DIR final adjoint  the compensated film curves: gradients of the log raw
//                   exposure and the filtered corrections, the film gamma
//                   and (through the compensated axes) the donor K;
//   DIR operator    the spatial DIR filter is symmetric (DCT-II diagonal on
//                   the half-sample-symmetric extension), so its adjoint is
//                   the same filter, run by the same launches;
//   source adjoint  Langmuir donors, film curves, log, exposure scale and the
//                   (piecewise linear) Mallett reconstruction to the input RGB.
// Parameter gradients are summed per block in a fixed order and reduced by
// one block, so the result is deterministic.
#include <hip/hip_runtime.h>

#include "fj_fit.h"
#include "fj_pixel_math.h"

namespace fjx {

namespace {

constexpr int kFitBlock = 256;
constexpr float kLn2 = 0.69314718055994530942f;
constexpr float kLn10 = 2.30258509299404568402f;

// ---- forward-mode values with three tangents ------------------------------
struct T3 {
    float v, a, b, c;
};
__device__ __forceinline__ T3 tc(float v) { return T3{v, 0.0f, 0.0f, 0.0f}; }
__device__ __forceinline__ T3 chain(const T3& x, float v, float dv) {
    return T3{v, dv * x.a, dv * x.b, dv * x.c};
}
__device__ __forceinline__ T3 operator+(const T3& x, const T3& y) {
    return T3{x.v + y.v, x.a + y.a, x.b + y.b, x.c + y.c};
}
__device__ __forceinline__ T3 operator-(const T3& x, const T3& y) {
    return T3{x.v - y.v, x.a - y.a, x.b - y.b, x.c - y.c};
}
__device__ __forceinline__ T3 operator+(const T3& x, float s) { return T3{x.v + s, x.a, x.b, x.c}; }
__device__ __forceinline__ T3 operator-(const T3& x, float s) { return T3{x.v - s, x.a, x.b, x.c}; }
__device__ __forceinline__ T3 operator*(float s, const T3& x) { return T3{s * x.v, s * x.a, s * x.b, s * x.c}; }
__device__ __forceinline__ T3 operator*(const T3& x, float s) { return s * x; }
__device__ __forceinline__ T3 operator*(const T3& x, const T3& y) {
    return T3{x.v * y.v, fmaf(x.a, y.v, x.v * y.a), fmaf(x.b, y.v, x.v * y.b), fmaf(x.c, y.v, x.v * y.c)};
}
__device__ __forceinline__ T3 operator/(const T3& x, float s) {
    const float r = 1.0f / s;
    return T3{x.v / s, x.a * r, x.b * r, x.c * r};
}
__device__ __forceinline__ T3 operator/(const T3& x, const T3& y) {
    const float q = x.v / y.v, r = 1.0f / y.v;
    return T3{q, (x.a - q * y.a) * r, (x.b - q * y.b) * r, (x.c - q * y.c) * r};
}
// a * b + c with the forward's fmaf for the value.
__device__ __forceinline__ T3 tfma(const T3& a, const T3& b, const T3& c) {
    return T3{fmaf(a.v, b.v, c.v), fmaf(a.a, b.v, fmaf(a.v, b.a, c.a)), fmaf(a.b, b.v, fmaf(a.v, b.b, c.b)),
              fmaf(a.c, b.v, fmaf(a.v, b.c, c.c))};
}
__device__ __forceinline__ T3 tfma(float a, const T3& b, const T3& c) {
    return T3{fmaf(a, b.v, c.v), fmaf(a, b.a, c.a), fmaf(a, b.b, c.b), fmaf(a, b.c, c.c)};
}
__device__ __forceinline__ T3 tfma(const T3& a, float b, float c) {
    return T3{fmaf(a.v, b, c), a.a * b, a.b * b, a.c * b};
}
// fminf(hi, fmaxf(lo, x)): the tangent survives only strictly inside.
__device__ __forceinline__ T3 tclamp(const T3& x, float lo, float hi) {
    const float v = fminf(hi, fmaxf(lo, x.v));
    return (x.v > lo && x.v < hi) ? T3{v, x.a, x.b, x.c} : tc(v);
}
__device__ __forceinline__ void tmat3v(const float* m, const T3* v, T3* o) {
    o[0] = tfma(m[0], v[0], tfma(m[1], v[1], m[2] * v[2]));
    o[1] = tfma(m[3], v[0], tfma(m[4], v[1], m[5] * v[2]));
    o[2] = tfma(m[6], v[0], tfma(m[7], v[1], m[8] * v[2]));
}
__device__ __forceinline__ T3 texp2(const T3& x) {
    const float y = hw_exp2(x.v);
    return chain(x, y, y * kLn2);
}
__device__ __forceinline__ T3 tcbrt(const T3& x) {
    const float y = scbrt(x.v);
    return chain(x, y, y != 0.0f ? 1.0f / (3.0f * y * y) : 0.0f);
}
// x^p for x > 0.
__device__ __forceinline__ T3 tpow(const T3& x, float p) {
    const float y = powf(x.v, p);
    return chain(x, y, x.v > 0.0f ? p * y / x.v : 0.0f);
}
__device__ __forceinline__ T3 thypot(const T3& x, const T3& y) {
    const float r = hypotf(x.v, y.v);
    const float ir = r > 0.0f ? 1.0f / r : 0.0f;
    return T3{r, (x.v * x.a + y.v * y.a) * ir, (x.v * x.b + y.v * y.b) * ir, (x.v * x.c + y.v * y.c) * ir};
}
__device__ __forceinline__ T3 tatan2(const T3& y, const T3& x) {
    const float h = atan2f(y.v, x.v);
    const float d = fmaf(x.v, x.v, y.v * y.v);
    const float id = d > 0.0f ? 1.0f / d : 0.0f;
    return T3{h, (x.v * y.a - y.v * x.a) * id, (x.v * y.b - y.v * x.b) * id, (x.v * y.c - y.v * x.c) * id};
}

// ---- the scanner tail with tangents (fj_pixel_math.h's scan_colour and
// encode_output, same operation order for the values) ------------------------
__device__ __forceinline__ T3 thermite(const T3& y0, const T3& y1, const T3& m0, const T3& m1, const T3& t) {
    const T3 t2 = t * t, t3 = t2 * t;
    const T3 h00 = tfma(2.0f, t3, tfma(-3.0f, t2, tc(1.0f)));
    const T3 h10 = tfma(-2.0f, t2, t3 + t);
    const T3 h01 = tfma(-2.0f, t3, 3.0f * t2);
    const T3 h11 = t3 - t2;
    return tfma(h00, y0, tfma(h10, m0, tfma(h01, y1, h11 * m1)));
}

__device__ __forceinline__ void tpchip_coord(const T3& v, int res, int& b, T3& t) {
    const float top = static_cast<float>(res - 1);
    const T3 c = tclamp(v * top, 0.0f, top);
    if (c.v >= top) {
        b = res - 2;
        t = tc(1.0f);
    } else {
        b = static_cast<int>(floorf(c.v));
        t = c - static_cast<float>(b);
    }
}

__device__ __forceinline__ bool tscan_lut(const ScanParams& sc, const T3 nrm[3], T3 out[3]) {
    if (!isfinite(nrm[0].v) || !isfinite(nrm[1].v) || !isfinite(nrm[2].v)) return false;
    const int r = sc.res;
    int c, m, y;
    T3 tcv, tmv, tyv;
    tpchip_coord(nrm[0], r, c, tcv);
    tpchip_coord(nrm[1], r, m, tmv);
    tpchip_coord(nrm[2], r, y, tyv);
    const int base = (c * r + m) * r + y;
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
        const T3 v00 = thermite(tc(n000.v[o]), tc(n100.v[o]), tc(n000.sc[o]), tc(n100.sc[o]), tcv);
        const T3 v10 = thermite(tc(n010.v[o]), tc(n110.v[o]), tc(n010.sc[o]), tc(n110.sc[o]), tcv);
        const T3 v01 = thermite(tc(n001.v[o]), tc(n101.v[o]), tc(n001.sc[o]), tc(n101.sc[o]), tcv);
        const T3 v11 = thermite(tc(n011.v[o]), tc(n111.v[o]), tc(n011.sc[o]), tc(n111.sc[o]), tcv);
        const T3 sm00 = tfma(tcv, n100.sm[o] - n000.sm[o], n000.sm[o]);
        const T3 sm10 = tfma(tcv, n110.sm[o] - n010.sm[o], n010.sm[o]);
        const T3 sm01 = tfma(tcv, n101.sm[o] - n001.sm[o], n001.sm[o]);
        const T3 sm11 = tfma(tcv, n111.sm[o] - n011.sm[o], n011.sm[o]);
        const T3 vy0 = thermite(v00, v10, sm00, sm10, tmv);
        const T3 vy1 = thermite(v01, v11, sm01, sm11, tmv);
        const T3 a0 = tfma(tcv, n100.sy[o] - n000.sy[o], n000.sy[o]);
        const T3 a1 = tfma(tcv, n110.sy[o] - n010.sy[o], n010.sy[o]);
        const T3 b0 = tfma(tcv, n101.sy[o] - n001.sy[o], n001.sy[o]);
        const T3 b1 = tfma(tcv, n111.sy[o] - n011.sy[o], n011.sy[o]);
        const T3 sy0 = tfma(tmv, a1 - a0, a0);
        const T3 sy1 = tfma(tmv, b1 - b0, b0);
        const T3 val = thermite(vy0, vy1, sy0, sy1, tyv);
        out[o] = tclamp(val, lo[o], hi[o]);
    }
    return true;
}

__device__ __forceinline__ T3 tknee(const T3& v, const float k[3]) {
    if (!(v.v > k[0])) return v;
    const float span = k[1] - k[0];
    const T3 nrm = (v - k[0]) / span;
    return (span * nrm) / tpow(tpow(nrm, k[2]) + 1.0f, 1.0f / k[2]) + k[0];
}

__device__ __forceinline__ T3 tgamut_boundary(const float* cmax, const T3& L, const T3& hue) {
    const float kPi = 3.14159265358979323846f, kTwoPi = 2.0f * kPi;
    const T3 cl = tclamp(L, 0.02f, 1.0f);
    const T3 lp = (cl - 0.02f) / (1.0f - 0.02f) * static_cast<float>(kGamutL - 1);
    const int l0 = static_cast<int>(floorf(lp.v));
    const int l1 = l0 + 1 < kGamutL ? l0 + 1 : l0;
    const T3 lf = lp - static_cast<float>(l0);
    float wh = fmodf(hue.v + kPi, kTwoPi);
    if (wh < 0.0f) wh += kTwoPi;
    const T3 hp = T3{wh, hue.a, hue.b, hue.c} / kTwoPi * static_cast<float>(kGamutH);
    const float hfl = floorf(hp.v);
    const int h0 = static_cast<int>(hfl) % kGamutH;
    const int h1 = (h0 + 1) % kGamutH;
    const T3 hf = hp - hfl;
    const float a0 = __ldg(cmax + l0 * kGamutH + h0), a1 = __ldg(cmax + l0 * kGamutH + h1);
    const float b0 = __ldg(cmax + l1 * kGamutH + h0), b1 = __ldg(cmax + l1 * kGamutH + h1);
    const T3 first = tfma(hf, a1 - a0, a0);
    const T3 second = tfma(hf, b1 - b0, b0);
    return tfma(lf, second - first, first);
}

__device__ __forceinline__ void tgamut_compress(const ColorParams& co, T3 rgb[3]) {
    const float kXyzLms[9] = {0.8189330101f, 0.3618667424f, -0.1288597137f, 0.0329845436f, 0.9293118715f,
                              0.0361456387f, 0.0482003018f, 0.2643662691f, 0.6338517070f};
    const float kLmsXyz[9] = {1.2270138511f, -0.5577999807f, 0.2812561490f, -0.0405801784f, 1.1122568696f,
                              -0.0716766787f, -0.0763812845f, -0.4214819784f, 1.5861632204f};
    const float kLmsLab[9] = {0.2104542553f, 0.7936177850f, -0.0040720468f, 1.9779984951f, -2.4285922050f,
                              0.4505937099f, 0.0259040371f, 0.7827717662f, -0.8086757660f};
    const float kLabLms[9] = {1.0f, 0.3963377774f, 0.2158037573f, 1.0f, -0.1055613458f,
                              -0.0638541728f, 1.0f, -0.0894841775f, -1.2914855480f};
    T3 xyz[3], lms[3], lab[3];
    tmat3v(co.native_to_xyz, rgb, xyz);
    tmat3v(kXyzLms, xyz, lms);
    lms[0] = tcbrt(lms[0]);
    lms[1] = tcbrt(lms[1]);
    lms[2] = tcbrt(lms[2]);
    tmat3v(kLmsLab, lms, lab);
    lab[0] = tknee(lab[0], co.lknee);
    const T3 chroma = thypot(lab[1], lab[2]);
    if (chroma.v > 0.0f) {
        const T3 hue = tatan2(lab[2], lab[1]);
        T3 bnd = tgamut_boundary(co.cmax, lab[0], hue);
        if (!(bnd.v > 1.0e-9f)) bnd = tc(1.0e-9f);
        const T3 comp = tknee(chroma / bnd, co.cknee) * bnd;
        const T3 s = comp / chroma;
        lab[1] = lab[1] * s;
        lab[2] = lab[2] * s;
    }
    tmat3v(kLabLms, lab, lms);
    lms[0] = lms[0] * lms[0] * lms[0];
    lms[1] = lms[1] * lms[1] * lms[1];
    lms[2] = lms[2] * lms[2] * lms[2];
    tmat3v(kLmsXyz, lms, xyz);
    tmat3v(co.xyz_to_native, xyz, rgb);
}

// Scanner normalized densities -> linear output (no glare, no CCTF): false on
// a scan error (the forward writes 0 there).
__device__ __forceinline__ bool tscan_output(const ScanParams& sc, const ColorParams& co, const T3 nrm[3],
                                             T3 out[3]) {
    T3 l2[3];
    if (!tscan_lut(sc, nrm, l2)) return false;
    T3 xyz[3] = {texp2(l2[0]), texp2(l2[1]), texp2(l2[2])};
    if (sc.correction) {
        const T3 cy = tclamp(tfma(xyz[1], sc.corr_slope, sc.corr_offset), 0.0f, 1.0f);
        const T3 s = cy / (xyz[1] + 1e-10f);
        xyz[0] = xyz[0] * s;
        xyz[1] = xyz[1] * s;
        xyz[2] = xyz[2] * s;
    }
    T3 rgb[3];
    tmat3v(sc.xyz_to_rgb, xyz, rgb);
    if (!isfinite(rgb[0].v) || !isfinite(rgb[1].v) || !isfinite(rgb[2].v)) return false;
    if (co.gamut) {
        tgamut_compress(co, rgb);
        if (!isfinite(rgb[0].v) || !isfinite(rgb[1].v) || !isfinite(rgb[2].v)) return false;
    }
    T3 lin[3];
    if (co.matrix) {
        tmat3v(co.to_output, rgb, lin);
    } else {
        lin[0] = rgb[0];
        lin[1] = rgb[1];
        lin[2] = rgb[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = tclamp(lin[c], 0.0f, 1.0f);
    return true;
}

// curve_at (fj_pixel_math.h) with its slope d value / d xq and segment (seg
// -1 where the value is clamped or the segment is degenerate: slope 0).
__device__ __forceinline__ float curve_slope(const float* x, const float* y, int n, float xq, float& slope,
                                             int& seg) {
    slope = 0.0f;
    seg = -1;
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
    slope = (y1 - y0) / den;
    seg = i;
    return fmaf((xq - x0) / den, y1 - y0, y0);
}

// Block sum of `kSlots` per-thread partials into partials[block][kFitSlots]
// at the slots `slot_of`, in a fixed order (deterministic).
template <int kSlots>
__device__ __forceinline__ void block_partials(float (&v)[kSlots], const int (&slot_of)[kSlots], float* out) {
    __shared__ float red[kSlots][kFitBlock];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) red[s][threadIdx.x] = v[s];
    __syncthreads();
    for (int w = kFitBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int s = 0; s < kSlots; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kFitSlots) {
        float r = 0.0f;
#pragma unroll
        for (int s = 0; s < kSlots; ++s)
            if (slot_of[s] == static_cast<int>(threadIdx.x)) r = red[s][0];
        out[threadIdx.x] = r;
    }
}

__device__ __forceinline__ float* stage_curves(const float* src, int count, float* lds) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) lds[i] = src[i];
    return lds;
}

}  // namespace

// Forward tail: grained density -> print (gamma) -> scan -> linear RGBA.
__global__ __launch_bounds__(kFitBlock) void fit_print_kernel(FitArgs a) {
    extern __shared__ float lds[];
    const PixelParams& p = a.pp;
    const bool in_lds = p.pr.n <= kMaxLdsCurve;
    const float* curves = in_lds ? stage_curves(p.pr.curves, 4 * p.pr.n, lds) : p.pr.curves;
    __syncthreads();
    const UniformF4 spec = uniform_table(p.pr.spec);
    const long long total = static_cast<long long>(p.io.width) * p.io.height;
    float* dens = a.saved + 6 * total;
    const int n = p.pr.n;
    unsigned int errors = 0u;
    for (long long idx = static_cast<long long>(blockIdx.x) * kFitBlock + threadIdx.x; idx < total;
         idx += static_cast<long long>(gridDim.x) * kFitBlock) {
        float D[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D[c] = dens[c * total + idx];
            if (a.delta) D[c] += a.delta[c * total + idx];
            dens[c * total + idx] = D[c];
        }
        float inner[3];
        print_inner(p.pr, spec, D, inner);
        float Dp[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float l = log10_fast(inner[c] * p.pr.outer_scale + 1e-10f);
            Dp[c] = sample_logE(curves, curves + (1 + c) * n, n, l, a.print_gamma);
        }
        float rgb[3];
        const bool ok = scan_colour(p.sc, p.co, Dp, 0.0f, rgb);
        if (ok) {
            encode_output(p.co, rgb);
        } else {
            errors |= 1u;
            rgb[0] = rgb[1] = rgb[2] = 0.0f;
        }
        const float4 s = reinterpret_cast<const float4*>(a.src)[idx];
        reinterpret_cast<float4*>(a.dst)[idx] = make_float4(rgb[0], rgb[1], rgb[2], s.w);
    }
    if (errors && p.error_word) atomicOr(p.error_word, errors);
}

// Print adjoint: loss gradient -> grained-density gradient; print exposure,
// Y/M filtration and print gamma partials.
__global__ __launch_bounds__(kFitBlock) void fit_print_adjoint_kernel(FitArgs a) {
    extern __shared__ float lds[];
    const PixelParams& p = a.pp;
    const bool in_lds = p.pr.n <= kMaxLdsCurve;
    const float* curves = in_lds ? stage_curves(p.pr.curves, 4 * p.pr.n, lds) : p.pr.curves;
    __syncthreads();
    const UniformF4 spec = uniform_table(p.pr.spec);
    const UniformF4 alpha = uniform_table(a.alpha);
    const long long total = static_cast<long long>(p.io.width) * p.io.height;
    const float* dens = a.saved + 6 * total;
    const int n = p.pr.n;
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // print exposure, Y, M, print gamma
    for (long long idx = static_cast<long long>(blockIdx.x) * kFitBlock + threadIdx.x; idx < total;
         idx += static_cast<long long>(gridDim.x) * kFitBlock) {
        const float4 g4 = reinterpret_cast<const float4*>(a.grad_dst)[idx];
        const float gout[3] = {g4.x, g4.y, g4.z};
        float D[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) D[c] = dens[c * total + idx];
        float inner[3];
        print_inner(p.pr, spec, D, inner);
        float v[3], l[3], sl[3];
        T3 nrm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = inner[c] * p.pr.outer_scale + 1e-10f;
            l[c] = log10_fast(v[c]);
            int seg;
            const float Dp = isfinite(l[c]) ? curve_slope(curves, curves + (1 + c) * n, n, l[c] * a.print_gamma, sl[c], seg)
                                            : __builtin_nanf("");
            if (!isfinite(l[c])) sl[c] = 0.0f;
            nrm[c] = T3{(Dp - p.sc.dmin[c]) * p.sc.inv_span[c], c == 0 ? 1.0f : 0.0f, c == 1 ? 1.0f : 0.0f,
                        c == 2 ? 1.0f : 0.0f};
        }
        T3 out[3];
        float gn[3] = {0.0f, 0.0f, 0.0f};
        if (tscan_output(p.sc, p.co, nrm, out)) {
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                gn[0] = fmaf(gout[o], out[o].a, gn[0]);
                gn[1] = fmaf(gout[o], out[o].b, gn[1]);
                gn[2] = fmaf(gout[o], out[o].c, gn[2]);
            }
        }
        float ga[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float gq = gn[c] * p.sc.inv_span[c] * sl[c];                  // d/d(gamma * l)
            part[3] = fmaf(gq, l[c], part[3]);                                  // print gamma
            const float gv = gq * a.print_gamma / (v[c] * kLn10);               // d/d(inner * scale + 1e-10)
            part[0] = fmaf(gv, inner[c] * a.print_route_scale, part[0]);       // print exposure
            ga[c] = gv * p.pr.outer_scale * p.pr.normalizer;                   // d/d(spectral sum)
            if (!isfinite(ga[c])) ga[c] = 0.0f;
        }
        // Spectral sum adjoint, recomputing each wavelength's transmittance.
        float gD0 = 0.0f, gD1 = 0.0f, gD2 = 0.0f;
#pragma unroll 4
        for (int i = 0; i < p.pr.k; ++i) {
            const float4 e = uniform_at(spec, 2 * i);
            const float4 s = uniform_at(spec, 2 * i + 1);
            const float4 al = uniform_at(alpha, i);
            const float ex = fmaf(D[0], e.x, fmaf(D[1], e.y, fmaf(D[2], e.z, e.w)));
            float t = hw_exp2(ex);
            t = (t == t) ? t : 0.0f;
            const float u = t * fmaf(ga[0], s.x, fmaf(ga[1], s.y, ga[2] * s.z));
            gD0 = fmaf(u, e.x, gD0);
            gD1 = fmaf(u, e.y, gD1);
            gD2 = fmaf(u, e.z, gD2);
            part[2] = fmaf(u, al.x, part[2]);  // M
            part[1] = fmaf(u, al.y, part[1]);  // Y
        }
        a.g_dens[idx] = gD0 * kLn2;
        a.g_dens[total + idx] = gD1 * kLn2;
        a.g_dens[2 * total + idx] = gD2 * kLn2;
    }
    const int slots[4] = {kFitPrintExposure, kFitFilterY, kFitFilterM, kFitPrintGamma};
    block_partials<4>(part, slots, a.partials + static_cast<long long>(blockIdx.x) * kFitSlots);
}

// DIR final adjoint: density gradient -> filtered-correction and log-raw
// gradients; film gamma and donor K (compensated axes) partials.
__global__ __launch_bounds__(kFitBlock) void fit_dir_final_adjoint_kernel(FitArgs a) {
    extern __shared__ float lds[];
    const PixelParams& p = a.pp;
    const int n = p.dev.n;
    const bool in_lds = n <= kMaxLdsCurve;
    float* film = in_lds ? stage_curves(p.dev.curves, 7 * n, lds) : nullptr;
    if (in_lds)
        for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) lds[7 * n + i] = a.daxis[i];
    __syncthreads();
    const float* curves = in_lds ? film : p.dev.curves;
    const float* daxis = in_lds ? lds + 7 * n : a.daxis;
    const long long total = static_cast<long long>(p.io.width) * p.io.height;
    const float* log_raw = a.saved;
    const float* filtered = a.saved + 3 * total;
    const float gamma = a.film_gamma;
    float part[2] = {0.0f, 0.0f};  // film gamma, donor K
    for (long long idx = static_cast<long long>(blockIdx.x) * kFitBlock + threadIdx.x; idx < total;
         idx += static_cast<long long>(gridDim.x) * kFitBlock) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float gD = a.g_dens[(2 - b) * total + idx];
            const float l = log_raw[b * total + idx] - filtered[b * total + idx];
            const float q = l * gamma;
            const float* cx = curves + (4 + b) * n;
            const float* cy = curves + (1 + b) * n;
            float slope;
            int seg;
            (void)curve_slope(cx, cy, n, q, slope, seg);
            float gl = gD * slope * gamma;
            float gk = 0.0f;
            if (seg >= 0) {
                const float x0 = cx[seg], x1 = cx[seg + 1], den = x1 - x0;
                const float dx0 = daxis[b * n + seg], dx1 = daxis[b * n + seg + 1];
                gk = gD * (cy[seg + 1] - cy[seg]) * ((q - x1) * dx0 - (q - x0) * dx1) / (den * den);
            }
            if (!isfinite(gl)) gl = 0.0f;
            part[0] = fmaf(gD * slope, l, part[0]);
            if (isfinite(gk)) part[1] += gk;
            a.g_filtered[b * total + idx] = -gl;
            a.g_lograw[b * total + idx] = gl;
        }
    }
    const int slots[2] = {kFitFilmGamma, kFitDirK};
    block_partials<2>(part, slots, a.partials + static_cast<long long>(blockIdx.x) * kFitSlots);
}

// Source adjoint: raw-correction and log-raw gradients -> input RGB
// gradient; exposure, film gamma and donor K partials.
__global__ __launch_bounds__(kFitBlock) void fit_source_adjoint_kernel(FitArgs a) {
    extern __shared__ float lds[];
    const PixelParams& p = a.pp;
    const int n = p.dev.n;
    const bool in_lds = n <= kMaxLdsCurve;
    const float* curves = in_lds ? stage_curves(p.dev.curves, 4 * n, lds) : p.dev.curves;
    __syncthreads();
    const long long total = static_cast<long long>(p.io.width) * p.io.height;
    const float gamma = a.film_gamma;
    const float mscale = p.ex.mallett_scale, scale = p.ex.scale;
    float part[3] = {0.0f, 0.0f, 0.0f};  // exposure, film gamma, donor K
    for (long long idx = static_cast<long long>(blockIdx.x) * kFitBlock + threadIdx.x; idx < total;
         idx += static_cast<long long>(gridDim.x) * kFitBlock) {
        const float4 in = reinterpret_cast<const float4*>(a.src)[idx];
        const float rgb[3] = {fin0(in.x), fin0(in.y), fin0(in.z)};
        float s[3], sp[3];
        mat3v(p.ex.rgb_to_srgb, rgb, s);
#pragma unroll
        for (int j = 0; j < 3; ++j) sp[j] = fmaxf(0.0f, fin0(s[j]));
        float ge[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float e = fmaf(a.mallett_A[3 * b], sp[0], fmaf(a.mallett_A[3 * b + 1], sp[1], a.mallett_A[3 * b + 2] * sp[2]));
            const float E = fmaxf(0.0f, fmaxf(0.0f, e) * mscale * scale);
            const float logE = log10_fast(E + 1e-10f);
            float slope;
            int seg;
            const float layer = curve_slope(curves, curves + (1 + b) * n, n, logE * gamma, slope, seg);
            // Langmuir donor v (k + ref) / (k + v) and its derivatives.
            const float k = p.dev.donor_k[b], ref = p.dev.dref[b];
            const float den = k + layer;
            const float dv = (k + ref) * k / (den * den);
            const float dk = layer * (layer - ref) / (den * den);
            float gdonor = 0.0f;
#pragma unroll
            for (int r = 0; r < 3; ++r) gdonor = fmaf(p.dev.M[3 * b + r], a.g_filtered[r * total + idx], gdonor);
            const float glayer = gdonor * dv;
            if (isfinite(dk)) part[2] = fmaf(gdonor * dk, a.donor_k_unit[b], part[2]);
            const float gq = glayer * slope;
            part[1] = fmaf(gq, logE, part[1]);
            const float glog = a.g_lograw[b * total + idx] + gq * gamma;
            const float gE = glog / ((E + 1e-10f) * kLn10);
            const bool live = e > 0.0f && E > 0.0f && isfinite(gE);
            if (live) part[0] = fmaf(gE, E * kLn2, part[0]);
            ge[b] = live ? gE * mscale * scale : 0.0f;
        }
        float gs[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float g = fmaf(ge[0], a.mallett_A[j], fmaf(ge[1], a.mallett_A[3 + j], ge[2] * a.mallett_A[6 + j]));
            gs[j] = s[j] > 0.0f ? g : 0.0f;
        }
        float gr[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float* m = p.ex.rgb_to_srgb;
            gr[i] = isfinite(in.x) && isfinite(in.y) && isfinite(in.z)
                        ? fmaf(m[i], gs[0], fmaf(m[3 + i], gs[1], m[6 + i] * gs[2]))
                        : 0.0f;
        }
        reinterpret_cast<float4*>(a.grad_src)[idx] = make_float4(gr[0], gr[1], gr[2], 0.0f);
    }
    const int slots[3] = {kFitEv, kFitFilmGamma, kFitDirK};
    block_partials<3>(part, slots, a.partials + static_cast<long long>(blockIdx.x) * kFitSlots);
}

// Sums `count` partial rows of kFitSlots floats in a fixed order into out[8]:
// 1024 threads, each summing every 128th row of its slot (eight independent
// loads in flight per step), then a fixed-order tree over the 128 groups.
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void fit_reduce_kernel(const float* partials, int count, float* out) {
    __shared__ float red[kReduceThreads];
    const int slot = threadIdx.x % kFitSlots, group = threadIdx.x / kFitSlots;
    constexpr int kGroups = kReduceThreads / kFitSlots;
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int r = group;
    for (; r + 7 * kGroups < count; r += 8 * kGroups) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += partials[static_cast<long long>(r + u * kGroups) * kFitSlots + slot];
    }
    for (; r < count; r += kGroups) acc[0] += partials[static_cast<long long>(r) * kFitSlots + slot];
    red[threadIdx.x] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    for (int w = kGroups / 2; w > 0; w >>= 1) {
        if (group < w) red[threadIdx.x] += red[threadIdx.x + w * kFitSlots];
        __syncthreads();
    }
    if (threadIdx.x < kFitSlots) out[threadIdx.x] = threadIdx.x < kFitUsed ? red[threadIdx.x] : 0.0f;
}

// ---- launchers --------------------------------------------------------------

int fit_blocks(long long total) {
    const long long tiles = (total + kFitBlock - 1) / kFitBlock;
    return static_cast<int>(tiles < 2048 ? tiles : 2048);
}

static size_t curve_lds(int floats) { return floats <= 0 ? 0 : static_cast<size_t>(floats) * sizeof(float); }

hipError_t launch_fit_print(const FitArgs& a, hipStream_t stream) {
    const long long total = static_cast<long long>(a.pp.io.width) * a.pp.io.height;
    const size_t lds = a.pp.pr.n <= kMaxLdsCurve ? curve_lds(4 * a.pp.pr.n) : 0;
    hipLaunchKernelGGL(fit_print_kernel, dim3(fit_blocks(total)), dim3(kFitBlock), lds, stream, a);
    return hipGetLastError();
}

// The three adjoint launches write their partial sums to regions 0, 1 and 2
// of a.partials (a.partial_blocks rows each).
hipError_t launch_fit_adjoint(const FitArgs& a, int stage, hipStream_t stream) {
    const long long total = static_cast<long long>(a.pp.io.width) * a.pp.io.height;
    const int blocks = fit_blocks(total);
    if (blocks != a.partial_blocks) return hipErrorInvalidValue;
    FitArgs b = a;
    b.partials = a.partials + static_cast<long long>(stage) * blocks * kFitSlots;
    const int n = a.pp.dev.n;
    switch (stage) {
        case 0: {
            const size_t lds = a.pp.pr.n <= kMaxLdsCurve ? curve_lds(4 * a.pp.pr.n) : 0;
            hipLaunchKernelGGL(fit_print_adjoint_kernel, dim3(blocks), dim3(kFitBlock), lds, stream, b);
            break;
        }
        case 1: {
            const size_t lds = n <= kMaxLdsCurve ? curve_lds(10 * n) : 0;
            hipLaunchKernelGGL(fit_dir_final_adjoint_kernel, dim3(blocks), dim3(kFitBlock), lds, stream, b);
            break;
        }
        case 2: {
            const size_t lds = n <= kMaxLdsCurve ? curve_lds(4 * n) : 0;
            hipLaunchKernelGGL(fit_source_adjoint_kernel, dim3(blocks), dim3(kFitBlock), lds, stream, b);
            break;
        }
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fit_reduce(const float* partials, int rows, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(fit_reduce_kernel, dim3(1), dim3(kReduceThreads), 0, stream, partials, rows, out);
    return hipGetLastError();
}

}  // namespace fjx
