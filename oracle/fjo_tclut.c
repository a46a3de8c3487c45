/* fjo_tclut.c — the oracle's restatement of the film TC-LUT build (test
 * infrastructure: loaded only by tests/).
 *
 * The TC LUT is a 192 x 192 table over the tri2quad chromaticity square of
 * the film's raw R/G/B response to the reconstructed spectra
 * (rust/film-juicer-core/src/reconstruction.rs:594-677): per cell, the 81
 * spectral samples (optionally blurred along wavelength, :540-569) weighted by
 * the sensitivities (Arctic: relit by the reference illuminant and divided by
 * the neutral response of the projected white, :609-626); Hanatos cells may
 * take the surface correction exp2(surface(tc)) (:570-593, :644-646); with
 * input compression the table is remapped through the input gamut hull
 * (gamut.rs:256-364). The reference builds it in FP32 (the table's words are
 * its contract), so this restatement keeps FP32 with the reference's order of
 * operations (compiled without contraction) and its FP64 hull geometry. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fj_oracle.h"

#define TC_CELLS (192 * 192)
#define SPECTRA (TC_CELLS * 81)

/* f32::clamp(0, 1): NaN stays NaN, -0 stays -0. */
static float clamp01(float v) { return isnan(v) ? v : (v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v)); }

static void tri2quad(float tx, float ty, float out[2]) {
    const float x = 1.0f - tx;
    const float den = x < 1e-10f ? 1e-10f : x;
    out[0] = clamp01(x * x);
    out[1] = clamp01(ty / den);
}

static int projected_white(const float w[3], float out[2]) {
    const float sum = (w[0] + w[1]) + w[2];
    if (!isfinite(sum) || sum <= 0.0f) return 0;
    tri2quad(w[0] / sum, w[1] / sum, out);
    return isfinite(out[0]) && isfinite(out[1]);
}

static float mitchell_weight(float t) {
    const float B = 1.0f / 3.0f, C = 1.0f / 3.0f;
    const float x = fabsf(t);
    if (x < 1.0f)
        return (1.0f / 6.0f) * ((12.0f - 9.0f * B - 6.0f * C) * x * x * x + (-18.0f + 12.0f * B + 6.0f * C) * x * x +
                                (6.0f - 2.0f * B));
    if (x < 2.0f)
        return (1.0f / 6.0f) * ((-B - 6.0f * C) * x * x * x + (6.0f * B + 30.0f * C) * x * x +
                                (-12.0f * B - 48.0f * C) * x + (8.0f * B + 24.0f * C));
    return 0.0f;
}

static int mitchell_coordinate(float normalized, int* base, float* frac) {
    const float v = clamp01(normalized) * 191.0f;
    if (!isfinite(v)) return 0;
    if (v >= 191.0f) {
        *base = 190;
        *frac = 1.0f;
        return 1;
    }
    *base = (int)floorf(v);
    *frac = v - (float)*base;
    return 1;
}

static int mitchell_reflect(int i) { return i < 0 ? -i : (i >= 192 ? 382 - i : i); }

static int sample_mitchell(const float* s, const float tc[2], int stride, int component, float* out) {
    int cb, mb;
    float cf, mf;
    if (!mitchell_coordinate(tc[0], &cb, &cf) || !mitchell_coordinate(tc[1], &mb, &mf)) return 0;
    const float cw[4] = {mitchell_weight(cf + 1.0f), mitchell_weight(cf), mitchell_weight(cf - 1.0f),
                         mitchell_weight(cf - 2.0f)};
    const float mw[4] = {mitchell_weight(mf + 1.0f), mitchell_weight(mf), mitchell_weight(mf - 1.0f),
                         mitchell_weight(mf - 2.0f)};
    float value = 0.0f, weight_sum = 0.0f;
    for (int dc = 0; dc < 4; ++dc) {
        const int c = mitchell_reflect(cb - 1 + dc);
        for (int dm = 0; dm < 4; ++dm) {
            const int m = mitchell_reflect(mb - 1 + dm);
            const float weight = cw[dc] * mw[dm];
            weight_sum += weight;
            value += weight * s[((size_t)c * 192 + m) * stride + component];
        }
    }
    *out = weight_sum != 0.0f ? value / weight_sum : 0.0f;
    return 1;
}

static float hanatos_surface(const float* p, const float tc[2], const float center[2]) {
    const float x = tc[0] - center[0], y = tc[1] - center[1];
    const float x2 = x * x, y2 = y * y, x3 = x2 * x, y3 = y2 * y;
    const float raw = p[1] * x + p[2] * y + p[3] * x2 + p[4] * y2 + p[5] * x * y + p[6] * x3 + p[7] * y3 +
                      p[8] * x2 * y + p[9] * x * y2 + p[10] * x2 * x2 + p[11] * y2 * y2 + p[12] * x3 * y +
                      p[13] * x2 * y2 + p[14] * x * y3;
    const float normalized = raw / 2.0f;
    return raw / sqrtf(1.0f + normalized * normalized);
}

/* Gaussian kernel of the spectral blur (reconstruction.rs:227-268). */
static int blur_kernel(float sigma, float** out, int* radius) {
    const float rounded = 4.0f * sigma + 0.5f;
    if (!isfinite(rounded) || rounded < 1.0f) return FJO_TC_INVALID_KERNEL;
    if ((double)rounded >= 9.2233720368547758e18) return FJO_TC_IMPOSSIBLE_KERNEL;
    const long r = (long)rounded;
    if (r > 1000000) return FJO_TC_IMPOSSIBLE_KERNEL;
    float* k = (float*)malloc(sizeof(float) * (size_t)(2 * r + 1));
    const float s2 = sigma * sigma;
    float sum = 0.0f;
    for (long i = 0; i <= 2 * r; ++i) {
        const long off = i - r;
        k[i] = expf(-0.5f * (float)(off * off) / s2);
        sum += k[i];
    }
    if (!isfinite(sum) || sum <= 0.0f) {
        free(k);
        return FJO_TC_INVALID_KERNEL;
    }
    for (long i = 0; i <= 2 * r; ++i) k[i] /= sum;
    *out = k;
    *radius = (int)r;
    return 0;
}

static int gaussian_reflect(long i) {
    long v = i % 162;
    if (v < 0) v += 162;
    return (int)(v >= 81 ? 161 - v : v);
}

/* Input gamut compression of one chromaticity (gamut.rs:256-321). */
static int compress(const float center_f[2], const float* hull_xy, const float xy[2], float out[2]) {
    if (!isfinite(xy[0]) || !isfinite(xy[1])) return 0;
    const double c0 = center_f[0], c1 = center_f[1];
    const double dx = (double)xy[0] - c0, dy = (double)xy[1] - c1;
    const double distance = hypot(dx, dy);
    if (distance < 1e-9) {
        out[0] = xy[0];
        out[1] = xy[1];
        return 1;
    }
    const double dir0 = dx / distance, dir1 = dy / distance;
    double minimum = INFINITY;
    for (int e = 0; e < 1024; ++e) {
        const double ax = hull_xy[2 * e], ay = hull_xy[2 * e + 1];
        const double ex = (double)hull_xy[2 * e + 2] - ax, ey = (double)hull_xy[2 * e + 3] - ay;
        const double den = dir0 * ey - dir1 * ex;
        if (fabs(den) <= 1e-12) continue;
        const double ox = c0 - ax, oy = c1 - ay;
        const double d = (-ox * ey + oy * ex) / den;
        if (d > 1e-9 && d < minimum) {
            const double pos = (-ox * dir1 + oy * dir0) / den;
            if (pos >= -1e-9 && pos <= 1.0 + 1e-9) minimum = d;
        }
    }
    if (!isfinite(minimum) || minimum <= 1e-12) return 0;
    double k = distance / minimum;
    if (k > 0.815) {
        const double scale = 1.0 - 0.815;
        const double n = (k - 0.815) / scale;
        k = 0.815 + scale * (n / pow(1.0 + pow(n, 1.2), 1.0 / 1.2));
    }
    const double cd = k * minimum;
    const double r0 = c0 + dir0 * cd, r1 = c1 + dir1 * cd;
    if (!isfinite(r0) || !isfinite(r1)) return 0;
    out[0] = (float)r0;
    out[1] = (float)r1;
    return 1;
}

static int bilinear(const float* s, const float tc[2], int ch, float* out) {
    const float cp = clamp01(tc[0]) * 191.0f, mp = clamp01(tc[1]) * 191.0f;
    if (!isfinite(cp) || !isfinite(mp)) return 0;
    const int c0 = (int)floorf(cp), m0 = (int)floorf(mp);
    const int c1 = c0 + 1 < 191 ? c0 + 1 : 191, m1 = m0 + 1 < 191 ? m0 + 1 : 191;
    const float cf = cp - (float)c0, mf = mp - (float)m0;
#define V(c, m) s[((size_t)(c) * 192 + (m)) * 4 + ch]
    const float row0 = V(c0, m0) + cf * (V(c1, m0) - V(c0, m0));
    const float row1 = V(c0, m1) + cf * (V(c1, m1) - V(c0, m1));
#undef V
    *out = row0 + mf * (row1 - row0);
    return 1;
}

int fjo_tc_lut_build(const FjoTcLutInput* in, float* out) {
    for (int i = 0; i < 81; ++i)
        if (!isfinite(in->reference_illuminant[i]) || !(in->reference_illuminant[i] >= 0.0f))
            return FJO_TC_INVALID_ILLUMINANT;
    const int arctic = in->method == 2;
    const float* selected = in->spectra;
    float* blurred = NULL;
    if (!arctic && in->spectral_blur > 0.0f) {
        const float sigma = in->spectral_blur;
        if (!isfinite(sigma) || sigma < 0.0f) return FJO_TC_INVALID_BLUR;
        float* k = NULL;
        int r = 0;
        const int rc = blur_kernel(sigma, &k, &r);
        if (rc) return rc;
        blurred = (float*)malloc(sizeof(float) * SPECTRA);
        for (int cell = 0; cell < TC_CELLS; ++cell) {
            const size_t base = (size_t)cell * 81;
            for (int s = 0; s < 81; ++s) {
                float v = 0.0f;
                for (int i = 0; i <= 2 * r; ++i) v += k[i] * in->spectra[base + gaussian_reflect((long)s + i - r)];
                if (!isfinite(v)) {
                    free(k);
                    free(blurred);
                    return FJO_TC_NONFINITE_BLURRED;
                }
                blurred[base + s] = v;
            }
        }
        free(k);
        selected = blurred;
    }
    float denom[3] = {1.0f, 1.0f, 1.0f};
    int rc = 0;
    if (arctic) {
        float tc[2];
        if (!projected_white(in->projection_white_xyz, tc)) rc = FJO_TC_INVALID_ARCTIC_WHITE;
        denom[0] = denom[1] = denom[2] = 0.0f;
        for (int s = 0; !rc && s < 81; ++s) {
            float reflectance;
            if (!sample_mitchell(selected, tc, 81, s, &reflectance)) {
                rc = FJO_TC_UNUSABLE_SAMPLING;
                break;
            }
            const float relit = reflectance * in->reference_illuminant[s];
            for (int c = 0; c < 3; ++c) denom[c] += relit * in->sensitivity_rgb[s * 3 + c];
        }
        if (!rc)
            for (int c = 0; c < 3; ++c)
                if (!isfinite(denom[c]) || !(denom[c] > 0.0f)) rc = FJO_TC_INVALID_NEUTRAL;
    }
    float center[2] = {0.0f, 0.0f};
    const int surface = !arctic && in->apply_surface;
    if (!rc && surface && !projected_white(in->projection_white_xyz, center)) rc = FJO_TC_INVALID_SURFACE_WHITE;
    float* table = out;
    float* integrated = in->compression_active ? (float*)calloc(TC_CELLS * 4, sizeof(float)) : out;
    if (!in->compression_active) memset(out, 0, sizeof(float) * TC_CELLS * 4);
    for (int c = 0; !rc && c < 192; ++c) {
        const float tc_c = (float)c / 191.0f;
        for (int m = 0; !rc && m < 192; ++m) {
            const float tc_m = (float)m / 191.0f;
            float raw[3] = {0.0f, 0.0f, 0.0f};
            const size_t base = ((size_t)c * 192 + m) * 81;
            for (int s = 0; s < 81; ++s) {
                float energy = selected[base + s];
                if (arctic) energy *= in->reference_illuminant[s];
                for (int ch = 0; ch < 3; ++ch) raw[ch] += energy * in->sensitivity_rgb[s * 3 + ch];
            }
            for (int ch = 0; ch < 3; ++ch) {
                float value = raw[ch] / denom[ch];
                if (surface) {
                    const float tc[2] = {tc_c, tc_m};
                    value *= exp2f(hanatos_surface(in->surface_rgb + 15 * ch, tc, center));
                }
                if (!isfinite(value)) {
                    rc = FJO_TC_NONFINITE_INTEGRATED;
                    break;
                }
                integrated[((size_t)c * 192 + m) * 4 + ch] = value;
            }
        }
    }
    free(blurred);
    if (!rc && in->compression_active) {
        if (!in->hull_available) {
            rc = FJO_TC_MISSING_HULL;
        } else {
            memset(table, 0, sizeof(float) * TC_CELLS * 4);
            for (int c = 0; !rc && c < 192; ++c) {
                const float root = sqrtf((float)c / 191.0f);
                for (int m = 0; !rc && m < 192; ++m) {
                    const float xy[2] = {1.0f - root, ((float)m / 191.0f) * root};
                    float cx[2], tc[2];
                    if (!compress(in->hull_center_xy, in->hull_xy, xy, cx)) {
                        rc = FJO_TC_UNUSABLE_COMPRESSED;
                        break;
                    }
                    tri2quad(cx[0], cx[1], tc);
                    for (int ch = 0; ch < 3; ++ch) {
                        float v;
                        if (!bilinear(integrated, tc, ch, &v)) {
                            rc = FJO_TC_UNUSABLE_SAMPLING;
                            break;
                        }
                        if (!isfinite(v)) {
                            rc = FJO_TC_NONFINITE_REMAPPED;
                            break;
                        }
                        table[((size_t)c * 192 + m) * 4 + ch] = v;
                    }
                }
            }
        }
    }
    if (in->compression_active) free(integrated);
    return rc;
}
