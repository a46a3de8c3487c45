/*
 * fj_oracle.c — FP64 CPU oracle of the per-frame render path.
 *
 * TEST INFRASTRUCTURE ONLY (see fj_oracle.h). Double-precision restatement of
 * the reference method; citations are relative to the reference repository.
 *
 * Stage order follows execute_print_schedule / execute_direct_schedule
 * (src/Cuda/JuicerCudaExecutor.cpp:2561-3141): camera exposure -> film
 * development (+DIR) -> [print exposure -> print development] -> scanner LUT
 * -> scanner colour -> output gamut -> [post effects] -> output encoding.
 * Every stage works on whole double planes so spatial operators can sit
 * between stages exactly where the reference schedule puts them. Camera auto
 * exposure (src/Cuda/JuicerCudaAutoExposure.cu) meters the frame first
 * (fjo_auto_exposure).
 */
#include "fj_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LOG2_10 3.32192809488736234787
#define NSPEC 81

/* ------------------------------------------------------------------------ */
/* scanner LUT                                                              */
/* ------------------------------------------------------------------------ */

/* scanner/lut.rs:108-138 (Fritsch–Carlson harmonic-mean PCHIP slopes with the
 * one-sided three-point endpoint rule). */
void fjo_pchip_slopes(const double* v, int n, double* out) {
    double d[512];
    double* deltas = n - 1 <= 512 ? d : (double*)malloc(sizeof(double) * (size_t)(n - 1));
    for (int i = 0; i < n - 1; ++i) deltas[i] = v[i + 1] - v[i];
    if (n == 2) {
        out[0] = out[1] = deltas[0];
    } else {
        for (int side = 0; side < 2; ++side) {
            const double a = side == 0 ? deltas[0] : deltas[n - 2];
            const double b = side == 0 ? deltas[1] : deltas[n - 3];
            const double s = 0.5 * (3.0 * a - b);
            double e;
            if (s * a <= 0.0) e = 0.0;
            else if (a * b < 0.0 && fabs(s) > fabs(3.0 * a)) e = 3.0 * a;
            else e = s;
            out[side == 0 ? 0 : n - 1] = e;
        }
        for (int i = 1; i < n - 1; ++i) {
            const double p = deltas[i - 1], q = deltas[i];
            out[i] = (p == 0.0 || q == 0.0 || p * q <= 0.0) ? 0.0 : 2.0 * p * q / (p + q);
        }
    }
    if (deltas != d) free(deltas);
}

/* scanner/lut.rs:140-183: density of a normalized LUT node, 81-λ
 * transmittance integration against illuminant-weighted CMFs, log10. */
static void lut_log_xyz(const float* ec, const float* em, const float* ey, const float* wx,
                        const float* wy, const float* wz, const float* base, int k, float inv_yn,
                        const float rmin[3], const float ispan[3], int medium, const double nrm[3],
                        double out[3]) {
    double dens[3];
    for (int c = 0; c < 3; ++c) {
        const double scaled = nrm[c] / (double)ispan[c];
        dens[c] = medium == 0 ? scaled - (double)rmin[c] : scaled + (double)rmin[c];
    }
    double xyz[3] = {0.0, 0.0, 0.0};
    const float* w[3] = {wx, wy, wz};
    for (int s = 0; s < k; ++s) {
        const double b = base ? (double)base[s] : 0.0;
        const double spectral = dens[0] * (double)ec[s] + dens[1] * (double)em[s] +
                                dens[2] * (double)ey[s] + b;
        const double t = pow(10.0, -spectral);
        for (int c = 0; c < 3; ++c) {
            const double weight = (double)w[c][s];
            if (isfinite(weight)) {
                const double prod = t * weight;
                if (!isnan(prod)) xyz[c] += prod;
            }
        }
    }
    for (int c = 0; c < 3; ++c) {
        const double sum = xyz[c] * (double)inv_yn;
        const double clamped = isnan(sum) ? 0.0 : (sum > 0.0 ? sum : 0.0);
        out[c] = log10(clamped + 1e-10);
    }
}

int fjo_scanner_lut_build(const float* ec, const float* em, const float* ey, const float* wx,
                          const float* wy, const float* wz, const float* base, int k,
                          float inv_yn, const float rmin[3], const float ispan[3], int medium,
                          int r, float* log2_xyz, float* sc, float* sm, float* sy, float* cmin,
                          float* cmax, uint64_t* fail_index, uint32_t* fail_channel) {
    if (r < 2 || !log2_xyz || !sc || !sm || !sy || !cmin || !cmax) return 2;
    const size_t count = (size_t)r * r * r * 3;
    double* nodes = (double*)malloc(sizeof(double) * count);
    if (!nodes) return 2;
#define IDX(c, m, y, ch) (((((size_t)(c) * r + (m)) * r + (y)) * 3) + (ch))
    for (int c = 0; c < r; ++c)
        for (int m = 0; m < r; ++m)
            for (int y = 0; y < r; ++y) {
                const double nrm[3] = {(double)c / (r - 1), (double)m / (r - 1), (double)y / (r - 1)};
                double lx[3] = {NAN, NAN, NAN};
                if (k > 0)
                    lut_log_xyz(ec, em, ey, wx, wy, wz, base, k, inv_yn, rmin, ispan, medium, nrm, lx);
                for (int ch = 0; ch < 3; ++ch) {
                    if (!isfinite(lx[ch])) {
                        if (fail_index) *fail_index = ((uint64_t)c * r + m) * r + y;
                        if (fail_channel) *fail_channel = (uint32_t)ch;
                        free(nodes);
                        return 1;
                    }
                    const size_t i = IDX(c, m, y, ch);
                    nodes[i] = lx[ch] * LOG2_10;
                    log2_xyz[i] = (float)nodes[i];
                }
            }
    double* line = (double*)malloc(sizeof(double) * (size_t)r * 2);
    double* slopes = line + r;
    float* outs[3] = {sc, sm, sy};
    for (int axis = 0; axis < 3; ++axis)
        for (int a = 0; a < r; ++a)
            for (int b = 0; b < r; ++b)
                for (int ch = 0; ch < 3; ++ch) {
                    for (int i = 0; i < r; ++i) {
                        const size_t p = axis == 0 ? IDX(i, a, b, ch) : axis == 1 ? IDX(a, i, b, ch) : IDX(a, b, i, ch);
                        line[i] = nodes[p];
                    }
                    fjo_pchip_slopes(line, r, slopes);
                    for (int i = 0; i < r; ++i) {
                        const size_t p = axis == 0 ? IDX(i, a, b, ch) : axis == 1 ? IDX(a, i, b, ch) : IDX(a, b, i, ch);
                        outs[axis][p] = (float)slopes[i];
                    }
                }
    free(line);
    const int s = r - 1;
    for (int c = 0; c < s; ++c)
        for (int m = 0; m < s; ++m)
            for (int y = 0; y < s; ++y)
                for (int ch = 0; ch < 3; ++ch) {
                    double lo = nodes[IDX(c, m, y, ch)], hi = lo;
                    for (int dc = 0; dc < 2; ++dc)
                        for (int dm = 0; dm < 2; ++dm)
                            for (int dy = 0; dy < 2; ++dy) {
                                const double v = nodes[IDX(c + dc, m + dm, y + dy, ch)];
                                if (v < lo) lo = v;
                                if (hi < v) hi = v;
                            }
                    const size_t i = ((((size_t)c * s + m) * s + y) * 3) + ch;
                    cmin[i] = (float)lo;
                    cmax[i] = (float)hi;
                }
#undef IDX
    free(nodes);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* plan                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct Curve {
    const float* x;
    const float* y; /* strided */
    int n;
    int stride;
} Curve;

struct FjoPlan {
    FjPreparedHostData p; /* shallow copy; spans must outlive the plan */
    int print;
    int res;
    float* lut; /* one allocation holding the six LUT arrays */
    float *l2, *sc, *sm, *sy, *cmin, *cmax;
    double manual_scale;
    FjoDiffusionFn diffusion; /* optional replacement of fjo_diffusion */
    void* diffusion_user;
    float* grain_capture[3]; /* optional copy of the densities entering the grain stage */
    /* The frame's glare stream inputs: floor(time in frames) and the source
     * image origin (JuicerCudaExecutor.cpp:2416-2420). */
    uint64_t frame_token;
    int origin_x, origin_y;
    /* Differentiable print route: the print curves read at print_gamma *
     * log E, and a fixed C/M/Y density increment in place of the grain stage
     * (NULL: none). */
    double print_gamma;
    const float* fit_delta[3];
};

static void set_message(char* msg, size_t cap, const char* text) {
    if (msg && cap) {
        snprintf(msg, cap, "%s", text);
    }
}

FjoPlan* fjo_plan_create(const FjPreparedHostData* src, char* msg, size_t cap) {
    if (!src) return NULL;
    if (src->film_exposure.method == FJ_RAW_MALLETT_2019
            ? (src->film_exposure.mallett_basis.count != 81u * 3u || src->film_exposure.mallett_illuminant.count != 81u)
            : src->film_exposure.tc_lut_rgba.count != 192u * 192u * 4u) {
        set_message(msg, cap, "oracle: TC LUT missing");
        return NULL;
    }
    FjoPlan* plan = (FjoPlan*)calloc(1, sizeof(FjoPlan));
    plan->p = *src;
    plan->print_gamma = 1.0;
    plan->print = src->route == FJ_ROUTE_NEGATIVE_PRINT || src->route == FJ_ROUTE_POSITIVE_PRINT;
    const int r = (int)src->scanner_lut.resolution;
    plan->res = r;
    const size_t nodes = (size_t)r * r * r * 3, cells = (size_t)(r - 1) * (r - 1) * (r - 1) * 3;
    plan->lut = (float*)malloc(sizeof(float) * (nodes * 4 + cells * 2));
    plan->l2 = plan->lut;
    plan->sc = plan->l2 + nodes;
    plan->sm = plan->sc + nodes;
    plan->sy = plan->sm + nodes;
    plan->cmin = plan->sy + nodes;
    plan->cmax = plan->cmin + cells;
    /* Medium and range follow JuicerCudaResources.cpp:3914-3931. */
    float rmin[3], ispan[3];
    for (int c = 0; c < 3; ++c) {
        rmin[c] = plan->print ? src->scanner_bounds.min_cmy[c] : -src->scanner_bounds.min_cmy[c];
        ispan[c] = src->scanner_bounds.inv_span_cmy[c];
    }
    const FjScannerSpectra* s = &src->scanner_spectra;
    uint64_t fi = 0;
    uint32_t fc = 0;
    const int rc = fjo_scanner_lut_build(s->dye_c.data, s->dye_m.data, s->dye_y.data, s->weighted_x.data,
                                         s->weighted_y.data, s->weighted_z.data,
                                         s->base_density.count ? s->base_density.data : NULL, NSPEC,
                                         s->inverse_y_normalization, rmin, ispan, plan->print ? 1 : 0, r,
                                         plan->l2, plan->sc, plan->sm, plan->sy, plan->cmin, plan->cmax, &fi, &fc);
    if (rc != 0) {
        set_message(msg, cap, "oracle: scanner LUT build failed");
        fjo_plan_destroy(plan);
        return NULL;
    }
    /* JuicerCudaFilmPayloads.cpp:102: the manual scale is exp2 of the EV. */
    plan->manual_scale = exp2((double)src->film_exposure.manual_exposure_ev);
    return plan;
}

void fjo_plan_destroy(FjoPlan* plan) {
    if (!plan) return;
    free(plan->lut);
    free(plan);
}

int fjo_plan_scanner_lut(const FjoPlan* plan, int* res, const float** l2, const float** sc,
                         const float** sm, const float** sy, const float** cmin, const float** cmax) {
    if (!plan) return 1;
    *res = plan->res;
    *l2 = plan->l2;
    *sc = plan->sc;
    *sm = plan->sm;
    *sy = plan->sy;
    *cmin = plan->cmin;
    *cmax = plan->cmax;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* per-pixel helpers                                                        */
/* ------------------------------------------------------------------------ */

static double fin0(double v) { return isfinite(v) ? v : 0.0; }

static void mat3(const float m[9], const double v[3], double o[3]) {
    o[0] = (double)m[0] * v[0] + (double)m[1] * v[1] + (double)m[2] * v[2];
    o[1] = (double)m[3] * v[0] + (double)m[4] * v[1] + (double)m[5] * v[2];
    o[2] = (double)m[6] * v[0] + (double)m[7] * v[1] + (double)m[8] * v[2];
}

/* JuicerCudaDeviceHelpers.cuh:179-225 (CCTF decode for BT.2020 / sRGB). */
static double decode_cctf(int space, double v) {
    const double x = fmax(0.0, fin0(v));
    if (space == FJ_INPUT_BT2020) {
        const double a = 1.09929681, th = 0.0812428791;
        return x < th ? x / 4.5 : pow((x + (a - 1.0)) / a, 1.0 / 0.45);
    }
    if (space == FJ_INPUT_SRGB_REC709) return x <= 0.04045 ? x / 12.92 : pow((x + 0.055) / 1.055, 2.4);
    return fin0(v);
}

/* JuicerCudaDeviceHelpers.cuh:266-282 (Mitchell–Netravali, B = C = 1/3). */
static double mitchell(double t) {
    const double B = 1.0 / 3.0, C = 1.0 / 3.0, x = fabs(t);
    if (x < 1.0)
        return ((12.0 - 9.0 * B - 6.0 * C) * x * x * x + (-18.0 + 12.0 * B + 6.0 * C) * x * x + (6.0 - 2.0 * B)) / 6.0;
    if (x < 2.0)
        return ((-B - 6.0 * C) * x * x * x + (6.0 * B + 30.0 * C) * x * x + (-12.0 * B - 48.0 * C) * x + (8.0 * B + 24.0 * C)) / 6.0;
    return 0.0;
}

static int reflect_idx(int i, int n) {
    if (n <= 1) return 0;
    if (i < 0) return -i;
    if (i >= n) return 2 * (n - 1) - i;
    return i;
}

/* JuicerCudaDeviceHelpers.cuh:284-327 (bicubic TC-LUT sample, normalized by
 * the weight sum). */
static void tc_sample(const float* lut, double qc, double qm, double out[3]) {
    const int N = 192;
    int base[2];
    double frac[2];
    const double q[2] = {qc, qm};
    for (int a = 0; a < 2; ++a) {
        const double coord = fmin(1.0, fmax(0.0, q[a])) * (N - 1);
        if (coord >= N - 1) {
            base[a] = N - 2;
            frac[a] = 1.0;
        } else {
            base[a] = (int)floor(coord);
            frac[a] = coord - base[a];
        }
    }
    double wc[4], wm[4];
    for (int i = 0; i < 4; ++i) {
        wc[i] = mitchell(frac[0] + 1.0 - i);
        wm[i] = mitchell(frac[1] + 1.0 - i);
    }
    double acc[3] = {0, 0, 0}, wsum = 0.0;
    for (int dc = 0; dc < 4; ++dc) {
        const int c = reflect_idx(base[0] - 1 + dc, N);
        for (int dm = 0; dm < 4; ++dm) {
            const int m = reflect_idx(base[1] - 1 + dm, N);
            const double w = wc[dc] * wm[dm];
            wsum += w;
            const float* t = lut + ((size_t)c * N + m) * 4;
            for (int ch = 0; ch < 3; ++ch) acc[ch] += w * (double)t[ch];
        }
    }
    for (int ch = 0; ch < 3; ++ch) out[ch] = wsum != 0.0 ? acc[ch] / wsum : 0.0;
}

/* JuicerCudaDeviceHelpers.cuh:34-89: clamped linear lookup on an ascending
 * axis after scaling logE by the channel gamma. */
static double curve_at(const Curve* cv, double logE, double gamma) {
    if (!isfinite(logE)) return NAN;
    const double xq = logE * gamma;
    const int n = cv->n;
#define CY(i) ((double)cv->y[(size_t)(i) * cv->stride])
    const double xmin = cv->x[0], xmax = cv->x[n - 1];
    if (!(xmax >= xmin)) return CY(0);
    if (xq <= xmin) return CY(0);
    if (xq >= xmax) return CY(n - 1);
    int lo = 1, hi = n - 1, i1 = n;
    while (lo <= hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (xq < (double)cv->x[mid]) {
            i1 = mid;
            hi = mid - 1;
        } else {
            lo = mid + 1;
        }
    }
    if (i1 > n - 1) return CY(n - 1);
    const double x0 = cv->x[i1 - 1], x1 = cv->x[i1], y0 = CY(i1 - 1), y1 = CY(i1);
    const double den = x1 - x0;
    if (!(den > 0.0) || !isfinite(den)) return y0;
    return y0 + (xq - x0) / den * (y1 - y0);
#undef CY
}

/* JuicerCudaDeviceHelpers.cuh:91-104. */
static double sanitize_inf_logE(double logE, const Curve* cv) {
    if (isfinite(logE) || isnan(logE)) return logE;
    const double xmin = cv->x[0], xmax = cv->x[cv->n - 1];
    if (!(xmax >= xmin)) return logE;
    return logE > 0.0 ? xmax : xmin;
}

/* JuicerCudaDeviceHelpers.cuh:1073-1123 (guarded Langmuir v·(k+ref)/(k+v)). */
static int langmuir(double v, double k, double ref, double* out) {
    *out = NAN;
    if (!isfinite(v) || !isfinite(k) || !(k > 0.0) || !isfinite(ref)) return 0;
    const double den = k + v;
    if (den == 0.0) return 0;
    *out = v * (k + ref) / den;
    return isfinite(*out);
}

/* ------------------------------------------------------------------------ */
/* stage kernels over planes                                                */
/* ------------------------------------------------------------------------ */

typedef struct Frame {
    const FjoPlan* plan;
    const float* src;
    int w, h, comps;
    ptrdiff_t src_stride;
    float* dst;
    ptrdiff_t dst_stride;
    double* plane[9]; /* scratch planes */
    double* carrier[3]; /* camera film linear exposure B/G/R, or NULL */
    double* plin[3];    /* enlarger print-linear exposure C/M/Y, or NULL */
    float* dens[3];     /* film density C/M/Y planes for the grain stage, or NULL */
    double* dens64[3];  /* the same in FP64 when a fixed increment replaces the grain */
    int dens_out;       /* the back end stores density instead of printing */
    double* linear;   /* [h][w][3] */
    uint32_t errors;
    double ae;          /* metered auto-exposure scale, 1 when off */
    double* glare;      /* scanner glare percent per pixel, or NULL */
    pthread_mutex_t lock;
} Frame;

static void tc_exposure(const float* lut, const double w[3], double E[3]);

/* Reconstructed film raw exposure before scaling, B/G/R order
 * (JuicerCudaDeviceHelpers.cuh:509-617, :619-703); the working XYZ and, for
 * Mallett, the clamped linear sRGB the spectral sum consumes are returned
 * when asked for. */
static void film_raw(const FjFilmExposure* fe, const float* px, double* xyz_out, double* srgb_out, double* sums_out,
                     double E[3]) {
    double rgb[3];
    const int decode = (fe->flags & FJ_FILM_DECODE_CCTF) != 0;
    for (int c = 0; c < 3; ++c) rgb[c] = decode ? decode_cctf((int)fe->input_color_space, px[c]) : fin0(px[c]);
    double xyz[3], w[3];
    mat3(fe->input_rgb_to_xyz, rgb, xyz);
    if (fe->flags & FJ_FILM_ADAPT_INPUT) {
        mat3(fe->input_xyz_adapt, xyz, w);
    } else {
        memcpy(w, xyz, sizeof(w));
    }
    if (fe->method == FJ_RAW_MALLETT_2019) {
        /* Mallett 2019 (JuicerCudaDeviceHelpers.cuh:495-507 input to linear
         * sRGB, :619-668 81-λ exposure, :699-703 green mid-gray scale). */
        double s[3];
        mat3(fe->xyz_to_linear_srgb, w, s);
        const double r = fmax(0.0, fin0(s[0])), g = fmax(0.0, fin0(s[1])), b = fmax(0.0, fin0(s[2]));
        if (xyz_out) memcpy(xyz_out, w, sizeof(w));
        if (srgb_out) {
            srgb_out[0] = r;
            srgb_out[1] = g;
            srgb_out[2] = b;
        }
        const float* basis = fe->mallett_basis.data;
        const float* ill = fe->mallett_illuminant.data;
        const float* sens = fe->sensitivity_rgb.data;
        double eb = 0.0, eg = 0.0, er = 0.0;
        for (int i = 0; i < NSPEC; ++i) {
            const double spd = (r * basis[i * 3] + g * basis[i * 3 + 1] + b * basis[i * 3 + 2]) * (double)ill[i];
            if (!isfinite(spd)) continue;
            if (isfinite(sens[i * 3 + 2])) eb += spd * sens[i * 3 + 2];
            if (isfinite(sens[i * 3 + 1])) eg += spd * sens[i * 3 + 1];
            if (isfinite(sens[i * 3 + 0])) er += spd * sens[i * 3 + 0];
        }
        if (sums_out) {
            sums_out[0] = eb;
            sums_out[1] = eg;
            sums_out[2] = er;
        }
        const double ms = fe->mallett_green_midgray_scale;
        E[0] = fmax(0.0, (isfinite(eb) ? fmax(0.0, eb) : 0.0) * ms);
        E[1] = fmax(0.0, (isfinite(eg) ? fmax(0.0, eg) : 0.0) * ms);
        E[2] = fmax(0.0, (isfinite(er) ? fmax(0.0, er) : 0.0) * ms);
    } else {
        for (int c = 0; c < 3; ++c) w[c] = fin0(w[c]);
        if (xyz_out) memcpy(xyz_out, w, sizeof(w));
        tc_exposure(fe->tc_lut_rgba.data, w, E);
    }
}

static void px_raw(const FjoPlan* pl, const float* px, double E[3]) {
    film_raw(&pl->p.film_exposure, px, NULL, NULL, NULL, E);
}

/* Film raw exposure with the auto-exposure (1 when off), manual and route
 * scales (compute_film_raw_device, JuicerCudaDeviceHelpers.cuh:709-763). */
static void px_exposure(const FjoPlan* pl, double ae, const float* px, double E[3]) {
    px_raw(pl, px, E);
    double scale = ae * pl->manual_scale * (double)pl->p.film_exposure.route_correction_scale;
    if (!(isfinite(scale) && scale > 0.0)) scale = 1.0;
    for (int c = 0; c < 3; ++c) E[c] = fmax(0.0, E[c] * scale);
}

/* Camera carrier exposure: auto-exposure and manual scales
 * (compute_camera_film_linear_exposure_device, JuicerCudaDeviceHelpers.cuh:
 * 709-744). */
static void px_camera(const FjoPlan* pl, double ae, const float* px, double E[3]) {
    px_raw(pl, px, E);
    double scale = ae * pl->manual_scale;
    if (!(isfinite(scale) && scale > 0.0)) scale = 1.0;
    for (int c = 0; c < 3; ++c) E[c] = fmax(0.0, E[c] * scale);
}

/* (e^t - t - 1) / t^2 (normalized_exponential, JuicerCudaHighlightBoost.cu:62-70). */
static double hb_norm(double t) {
    if (fabs(t) >= 1e-4) return (expm1(t) - t) / (t * t);
    return 0.5 + t * (1.0 / 6.0 + t * (1.0 / 24.0 + t / 120.0));
}

int fjo_highlight_boost(const FjHighlightBoost* hb, double* const planes[3], size_t n) {
    /* JuicerCudaHighlightBoost.cu:40-186: maximum m over the three planes
     * (any non-finite or negative sample: invalid carrier), bypass when
     * m <= protection, overflow when m * gain > FLT_MAX; samples s > p become
     * s + m (g - 1) f^2 N(k f) / N(k), f = (s - p) / (m - p), k = k0 (m - p) / m. */
    double m = 0.0;
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < n; ++i) {
            const double v = planes[c][i];
            if (!(isfinite(v) && v >= 0.0)) return 6;
            if (v > m) m = v;
        }
    const double p = hb->protection;
    if (m <= p) return 0;
    if (m * hb->peak_gain > (double)FLT_MAX) return 7;
    const double range = m - p, shape = hb->sharpness * range / m;
    const double inv = 1.0 / hb_norm(shape), delta = m * (hb->peak_gain - 1.0);
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < n; ++i) {
            const double v = planes[c][i];
            if (!(v > p)) continue;
            const double f = (v - p) / range;
            planes[c][i] = v + delta * f * f * hb_norm(shape * f) * inv;
        }
    return 0;
}

/* 2^(-log2 Y) with the reference's guards, rounded to the device float
 * (JuicerCudaAutoExposure.cu:500-521). */
static double ae_scale_from(double y) {
    double ev = 0.0;
    if (y > 0.0 && isfinite(y)) {
        const double c = -log(y) / log(2.0);
        if (isfinite(c)) ev = c;
    }
    const double s = exp2(ev);
    return (isfinite(s) && s > 0.0) ? (double)(float)s : 1.0;
}

/* Preview index -> source coordinate (JuicerCudaAutoExposure.cu:75-85). */
static int ae_coord(int i, int preview, int source) {
    if (preview <= 0 || source <= 0) return 0;
    const long long c = ((2LL * i + 1LL) * (long long)source) / (2LL * preview);
    return (int)(c < 0 ? 0 : (c > source - 1 ? source - 1 : c));
}

double fjo_auto_exposure(const FjoPlan* pl, const float* src, int width, int height, int comps, ptrdiff_t stride) {
    const FjAutoExposure* d = &pl->p.auto_exposure;
    if (!(pl->p.film_exposure.flags & FJ_FILM_AUTO_EXPOSURE)) return 1.0;
    const int w = d->preview_width, h = d->preview_height;
    const FjRect sb = d->source_bounds, mb = d->meter_bounds;
    if (w <= 0 || h <= 0 || sb.x1 != 0 || sb.y1 != 0 || sb.x2 != width || sb.y2 != height) return -1.0;
    const int method = (int)d->method;
    double sy[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
    float* means = (float*)malloc(sizeof(float) * (size_t)w * h);
    unsigned char* valid = (unsigned char*)calloc((size_t)w * h, 1);
    float max_y = 0.0f;
    const int maxd = w > h ? w : h;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int px = mb.x1 + ae_coord(x, w, mb.x2 - mb.x1), py = mb.y1 + ae_coord(y, h, mb.y2 - mb.y1);
            if (px < sb.x1 || px >= sb.x2 || py < sb.y1 || py >= sb.y2) continue;
            double E[3];
            px_raw(pl, src + (ptrdiff_t)(py - sb.y1) * stride + (ptrdiff_t)(px - sb.x1) * comps, E);
            const double mean = (E[0] + E[1] + E[2]) / 3.0;
            if (!isfinite(mean)) continue;
            means[(size_t)y * w + x] = (float)mean;
            valid[(size_t)y * w + x] = 1;
            if ((float)mean > max_y) max_y = (float)mean;
            /* metering_weights (JuicerCudaAutoExposure.cu:87-153) */
            double wt[4] = {0, 0, 0, 0};
            const double nx = ((double)x / w - 0.5) * w / maxd, ny = ((double)y / h - 0.5) * h / maxd;
            /* Ring and spot membership are decisions on the reference's FP32
             * radius (unfused single-precision operations, as the device meter
             * evaluates them). */
            const float inv_max = 1.0f / (float)maxd;
            const float fx = (((float)x / (float)w) - 0.5f) * (float)w * inv_max;
            const float fy = (((float)y / (float)h) - 0.5f) * (float)h * inv_max;
            const float fxx = fx * fx, fyy = fy * fy;
            const float radius = sqrtf(fxx + fyy);
            if (method == FJ_METER_AVERAGE || method == FJ_METER_MEDIAN) {
                wt[0] = 1.0;
            } else if (method == FJ_METER_CENTER_WEIGHTED) {
                wt[0] = exp(-(nx * nx) / (2.0 * 0.04)) * exp(-(ny * ny) / (2.0 * 0.04));
            } else if (method == FJ_METER_PARTIAL) {
                wt[0] = radius < 0.15f ? 1.0 : 0.0;
                wt[3] = 1.0;
            } else if (method == FJ_METER_MATRIX) {
                const int ch = h / 5, cw = w / 5;
                if (ch > 0 && cw > 0 && y / ch < 5 && x / cw < 5) {
                    const double dy = (y / ch - 2.0) / 2.0, dx = (x / cw - 2.0) / 2.0;
                    wt[0] = 0.5 * (1.0 + cos(3.14159265358979323846 * sqrt(dx * dx + dy * dy) / sqrt(2.0)));
                }
            } else if (method == FJ_METER_MULTI_ZONE) {
                if (radius < 0.05f) wt[0] = 1.0;
                else if (radius < 0.25f) wt[1] = 1.0;
                else if (radius < 0.50f) wt[2] = 1.0;
                wt[3] = 1.0;
            } else if (method == FJ_METER_HIGHLIGHT_WEIGHTED) {
                wt[0] = mean * mean;
                wt[3] = 1.0;
            }
            for (int l = 0; l < 4; ++l) {
                const double v = (isfinite(wt[l]) && wt[l] > 0.0) ? wt[l] : 0.0;
                sy[l] += mean * v;
                sw[l] += v;
            }
        }
    }
    double result = 1.0;
    if (method == FJ_METER_MEDIAN) {
        /* 2048-bin histogram median (:524-660) */
        if (max_y > 0.0f && isfinite(max_y)) {
            unsigned long long hist[2048] = {0}, total = 0;
            for (size_t i = 0; i < (size_t)w * h; ++i) {
                if (!valid[i]) continue;
                const float m = means[i] < 0.0f ? 0.0f : means[i];
                const float norm = fminf(1.0f, m / max_y);
                hist[(int)(norm * 2047.0f)]++;
                total++;
            }
            if (total) {
                const unsigned long long target = total / 2ull;
                unsigned long long cum = 0, in_bin = 0;
                int bin = 2047;
                for (int i = 0; i < 2048; ++i) {
                    if (cum + hist[i] > target) {
                        bin = i;
                        in_bin = hist[i];
                        break;
                    }
                    cum += hist[i];
                }
                const double frac = in_bin ? (double)(target - cum) / (double)in_bin : 0.0;
                result = ae_scale_from(((double)bin + frac) * ((double)max_y / 2048.0));
            }
        }
    } else {
        double y = 0.0;
        int has = 0;
        if (method == FJ_METER_PARTIAL) {
            const int l = sw[0] > 0.0 ? 0 : 3;
            if (sw[l] > 0.0) { y = sy[l] / sw[l]; has = 1; }
        } else if (method == FJ_METER_MULTI_ZONE) {
            const double ring[3] = {0.50, 0.30, 0.20};
            double sum = 0.0, tot = 0.0;
            for (int l = 0; l < 3; ++l)
                if (sw[l] > 0.0) { sum += ring[l] * sy[l] / sw[l]; tot += ring[l]; }
            if (tot > 0.0) { y = sum / tot; has = 1; }
            else if (sw[3] > 0.0) { y = sy[3] / sw[3]; has = 1; }
        } else if (method == FJ_METER_HIGHLIGHT_WEIGHTED && sw[0] < 1e-12) {
            if (sw[3] > 0.0) { y = sy[3] / sw[3]; has = 1; }
        } else if (sw[0] > 0.0) {
            y = sy[0] / sw[0];
            has = 1;
        }
        result = has ? ae_scale_from(y) : 1.0;
    }
    free(means);
    free(valid);
    return result;
}

/* Route correction applied when a carrier is developed
 * (apply_film_route_correction_device, JuicerCudaDeviceHelpers.cuh:746-755). */
static void carrier_exposure(const FjoPlan* pl, double* const carrier[3], size_t idx, double E[3]) {
    const double route = pl->p.film_exposure.route_correction_scale;
    for (int c = 0; c < 3; ++c) E[c] = fmax(0.0, carrier[c][idx] * route);
}

/* TC-LUT exposure (Hanatos / Arctic) from working XYZ. */
static void tc_exposure(const float* lut, const double w[3], double E[3]) {
    const double sum = w[0] + w[1] + w[2];
    if (!(fabs(sum) <= FLT_MAX)) {
        /* The FP32 brightness overflows: brightness x sample is not finite,
         * and a non-finite lane becomes 0 (reconstruction.rs:684-703). */
        E[0] = E[1] = E[2] = 0.0;
        return;
    }
    const double den = fmax(sum, 1e-10);
    const double x = w[0] / den, y = w[1] / den;
    /* tri2quad (JuicerCudaDeviceHelpers.cuh:244-252) */
    const double qx = fmin(1.0, fmax(0.0, (1.0 - x) * (1.0 - x)));
    const double qy = fmin(1.0, fmax(0.0, y / fmax(1.0 - x, 1e-10)));
    double rgbL[3];
    tc_sample(lut, qx, qy, rgbL);
    E[0] = fin0(sum * fin0(rgbL[2]));
    E[1] = fin0(sum * fin0(rgbL[1]));
    E[2] = fin0(sum * fin0(rgbL[0]));
}

static void film_curves(const FjoPlan* pl, Curve dens[3], Curve comp[3]) {
    const FjFilmDevelopment* fd = &pl->p.film_development;
    const int n = (int)fd->log_exposure.count;
    for (int b = 0; b < 3; ++b) {
        const int rgb = 2 - b; /* B/G/R device order reads R,G,B columns reversed */
        dens[b].x = fd->log_exposure.data;
        dens[b].y = fd->density_rgb.data + rgb;
        dens[b].n = n;
        dens[b].stride = 3;
        comp[b] = dens[b];
        if (pl->p.dir_couplers.mode != FJ_DIR_INACTIVE) comp[b].x = pl->p.dir_couplers.compensated_axes_rgb[rgb].data;
    }
}

/* Film development with non-spatial DIR: log raw -> layer densities ->
 * DIR corrections -> compensated curves (JuicerCudaScanPipeline.cu:1624-1697,
 * JuicerCudaDeviceHelpers.cuh:1125-1253). Output D in C/M/Y order. */
static uint32_t px_develop(const FjoPlan* pl, const double E[3], double Dcmy[3]) {
    const FjFilmDevelopment* fd = &pl->p.film_development;
    const FjDirCouplers* dir = &pl->p.dir_couplers;
    Curve dens[3], comp[3];
    film_curves(pl, dens, comp);
    const double gamma[3] = {fd->gamma_rgb[2], fd->gamma_rgb[1], fd->gamma_rgb[0]};
    double logE[3], layer[3];
    for (int b = 0; b < 3; ++b) {
        logE[b] = sanitize_inf_logE(log10(fmax(E[b], 0.0) + 1e-10), &dens[b]);
        layer[b] = curve_at(&dens[b], logE[b], gamma[b]);
    }
    if (dir->mode == FJ_DIR_INACTIVE) {
        Dcmy[0] = layer[2];
        Dcmy[1] = layer[1];
        Dcmy[2] = layer[0];
        return 0;
    }
    /* BGR packing of the DIR record (JuicerCudaFilmPayloads.cpp:172-190). */
    double M[9], dmax[3], dref[3], dk[3], cref[3], kr[3];
    for (int d = 0; d < 3; ++d) {
        for (int r = 0; r < 3; ++r) M[d * 3 + r] = dir->matrix_rgb[(2 - d) * 3 + (2 - r)];
        dmax[d] = dir->density_max_rgb[2 - d];
        dref[d] = dir->density_ref_rgb[2 - d];
        dk[d] = dir->donor_k_rgb[2 - d];
        cref[d] = dir->receiver_c_ref_rgb[2 - d];
        kr[d] = dir->receiver_kr_rgb[2 - d];
    }
    uint32_t fail = 0;
    double donors[3];
    for (int c = 0; c < 3; ++c) {
        donors[c] = dmax[c] - layer[c];
        if (dir->mode == FJ_DIR_NEGATIVE_DONOR_LANGMUIR && !langmuir(layer[c], dk[c], dref[c], &donors[c])) fail |= 1u << c;
    }
    double corr[3];
    for (int r = 0; r < 3; ++r) {
        corr[r] = M[0 * 3 + r] * donors[0] + M[1 * 3 + r] * donors[1] + M[2 * 3 + r] * donors[2];
        if (!isfinite(corr[r])) fail |= 1u << r;
    }
    if (dir->mode == FJ_DIR_POSITIVE_RECEIVER_LANGMUIR) {
        for (int c = 0; c < 3; ++c) {
            if (cref[c] == 0.0) {
                corr[c] = 0.0;
                continue;
            }
            double v;
            if (!langmuir(corr[c], kr[c], cref[c], &v)) fail |= 1u << c;
            corr[c] = v;
        }
    }
    if (fail == 0) {
        for (int c = 0; c < 3; ++c) {
            if (isnan(logE[c])) {
                fail |= 1u << c;
                continue;
            }
            const double o = logE[c];
            logE[c] = o - corr[c];
            if (!isfinite(logE[c]) && isfinite(o)) fail |= 1u << c;
        }
    }
    if (fail) {
        Dcmy[0] = Dcmy[1] = Dcmy[2] = NAN;
        return (1u << 8) | ((fail & 7u) << 9);
    }
    Dcmy[2] = curve_at(&comp[0], logE[0], gamma[0]);
    Dcmy[1] = curve_at(&comp[1], logE[1], gamma[1]);
    Dcmy[0] = curve_at(&comp[2], logE[2], gamma[2]);
    return 0;
}

/* Print exposure (JuicerCudaDeviceHelpers.cuh:866-964) and print development
 * (:1034-1071): 81-λ transmittance x filtered illuminant x paper sensitivity,
 * balance normalizer, preflash, print exposure, log10, print curves. */
/* 81-λ print exposure up to the mid-gray normalizer and preflash
 * (print_inner_raw_device, JuicerCudaDeviceHelpers.cuh:866-995). */
static void print_integral(const FjPrint* pr, const double D[3], double acc[3]) {
    const float* eps = pr->film_density_cmy.data;
    const float* base = pr->film_base_density.data;
    const float* sens = pr->sensitivity_cmy.data;
    const float* ill = pr->main_illuminant.data;
    acc[0] = acc[1] = acc[2] = 0.0;
    for (int i = 0; i < NSPEC; ++i) {
        const double s[3] = {sens[i * 3], sens[i * 3 + 1], sens[i * 3 + 2]};
        if (isnan(s[0]) && isnan(s[1]) && isnan(s[2])) continue;
        const double dl = D[0] * eps[i * 3] + D[1] * eps[i * 3 + 1] + D[2] * eps[i * 3 + 2] + (double)base[i];
        double e = pow(10.0, -dl) * (double)ill[i];
        if (isnan(e)) e = 0.0;
        for (int c = 0; c < 3; ++c)
            if (!isnan(s[c])) acc[c] += e * s[c];
    }
}

static void print_inner(const FjoPlan* pl, const double D[3], double acc[3]) {
    const FjPrint* pr = &pl->p.print;
    print_integral(pr, D, acc);
    const double pf = pr->preflash_exposure;
    const int preflash = (pr->flags & FJ_PRINT_PREFLASH) != 0;
    for (int c = 0; c < 3; ++c) {
        acc[c] *= (double)pr->normalizer;
        if (preflash && isfinite(pf) && pf > 0.0) acc[c] += (double)pr->preflash_raw_cmy[c] * pf;
    }
}

static void print_curves(const FjoPlan* pl, const double logp[3], double D[3]) {
    const FjPrint* pr = &pl->p.print;
    const int n = (int)pr->log_exposure.count;
    for (int c = 0; c < 3; ++c) {
        const Curve cv = {pr->log_exposure.data, pr->density_cmy.data + c, n, 3};
        D[c] = curve_at(&cv, logp[c], pl->print_gamma);
    }
}

/* Print exposure and development (apply_print_pipeline_device, :1047-1071). */
static void px_print(const FjoPlan* pl, double D[3]) {
    const FjPrint* pr = &pl->p.print;
    double acc[3], logp[3];
    print_inner(pl, D, acc);
    for (int c = 0; c < 3; ++c) {
        acc[c] *= (double)pr->exposure;
        acc[c] *= (double)pl->p.scanner_correction.exposure_scale;
        logp[c] = log10(acc[c] + 1e-10);
    }
    print_curves(pl, logp, D);
}

/* Enlarger-linear print exposure the enlarger diffusion blurs
 * (print_linear_for_diffusion_device, :997-1010), C/M/Y-forming planes. */
static void px_print_linear(const FjoPlan* pl, const double D[3], double lin[3]) {
    const FjPrint* pr = &pl->p.print;
    double acc[3];
    print_inner(pl, D, acc);
    const double outer = (double)(float)(pr->exposure * pl->p.scanner_correction.exposure_scale);
    for (int c = 0; c < 3; ++c) lin[c] = (fmax(acc[c], 0.0) + 1e-10) * outer;
}

/* Tricubic PCHIP sample of the log2 XYZ LUT (JuicerCudaDeviceHelpers.cuh:407-506). */
static double hermite(double y0, double y1, double m0, double m1, double t) {
    const double t2 = t * t, t3 = t2 * t;
    return (2 * t3 - 3 * t2 + 1) * y0 + (t3 - 2 * t2 + t) * m0 + (-2 * t3 + 3 * t2) * y1 + (t3 - t2) * m1;
}

static int px_scan_lut(const FjoPlan* pl, const double nrm[3], double out[3]) {
    const int r = pl->res;
    if (!isfinite(nrm[0]) || !isfinite(nrm[1]) || !isfinite(nrm[2])) return 0;
    int b[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
        double coord = fmin((double)(r - 1), fmax(0.0, nrm[a] * (r - 1)));
        if (coord >= r - 1) {
            b[a] = r - 2;
            t[a] = 1.0;
        } else {
            b[a] = (int)floor(coord);
            t[a] = coord - b[a];
        }
    }
#define N(arr, c, m, y, o) ((double)(arr)[((((size_t)(c) * r + (m)) * r + (y)) * 3) + (o)])
    const int c = b[0], m = b[1], y = b[2], s = r - 1;
    for (int o = 0; o < 3; ++o) {
        double v[2][2];
        for (int dm = 0; dm < 2; ++dm)
            for (int dy = 0; dy < 2; ++dy)
                v[dm][dy] = hermite(N(pl->l2, c, m + dm, y + dy, o), N(pl->l2, c + 1, m + dm, y + dy, o),
                                    N(pl->sc, c, m + dm, y + dy, o), N(pl->sc, c + 1, m + dm, y + dy, o), t[0]);
        double smv[2][2];
        for (int dm = 0; dm < 2; ++dm)
            for (int dy = 0; dy < 2; ++dy) {
                const double a0 = N(pl->sm, c, m + dm, y + dy, o), a1 = N(pl->sm, c + 1, m + dm, y + dy, o);
                smv[dm][dy] = a0 + t[0] * (a1 - a0);
            }
        const double vy0 = hermite(v[0][0], v[1][0], smv[0][0], smv[1][0], t[1]);
        const double vy1 = hermite(v[0][1], v[1][1], smv[0][1], smv[1][1], t[1]);
        double syv[2];
        for (int dy = 0; dy < 2; ++dy) {
            const double a00 = N(pl->sy, c, m, y + dy, o), a10 = N(pl->sy, c + 1, m, y + dy, o);
            const double a01 = N(pl->sy, c, m + 1, y + dy, o), a11 = N(pl->sy, c + 1, m + 1, y + dy, o);
            const double l0 = a00 + t[0] * (a10 - a00), l1 = a01 + t[0] * (a11 - a01);
            syv[dy] = l0 + t[1] * (l1 - l0);
        }
        const double val = hermite(vy0, vy1, syv[0], syv[1], t[2]);
        const size_t ci = ((((size_t)c * s + m) * s + y) * 3) + o;
        out[o] = fmin((double)pl->cmax[ci], fmax((double)pl->cmin[ci], val));
    }
#undef N
    return 1;
}

/* Output gamut compression in OkLch (src/GamutCompression.h:117-213). */
static const double OK_XYZ_LMS[9] = {0.8189330101, 0.3618667424, -0.1288597137, 0.0329845436, 0.9293118715,
                                     0.0361456387, 0.0482003018, 0.2643662691, 0.6338517070};
static const double OK_LMS_XYZ[9] = {1.2270138511, -0.5577999807, 0.2812561490, -0.0405801784, 1.1122568696,
                                     -0.0716766787, -0.0763812845, -0.4214819784, 1.5861632204};
static const double OK_LMSR_LAB[9] = {0.2104542553, 0.7936177850, -0.0040720468, 1.9779984951, -2.4285922050,
                                      0.4505937099, 0.0259040371, 0.7827717662, -0.8086757660};
static const double OK_LAB_LMSR[9] = {1.0, 0.3963377774, 0.2158037573, 1.0, -0.1055613458,
                                      -0.0638541728, 1.0, -0.0894841775, -1.2914855480};

static void mat3d(const double m[9], const double v[3], double o[3]) {
    o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

static double knee(double v, const float k[3]) {
    const double th = k[0], lim = k[1], pw = k[2];
    if (!(v > th)) return v;
    const double span = lim - th, n = (v - th) / span;
    return th + span * n / pow(1.0 + pow(n, pw), 1.0 / pw);
}

static double boundary(const float* cmax, double L, double hue) {
    const double pi = 3.14159265358979323846;
    const double cl = fmin(fmax(L, 0.02), 1.0);
    const double lp = (cl - 0.02) / (1.0 - 0.02) * 63.0;
    const int l0 = (int)floor(lp), l1 = l0 + 1 < 64 ? l0 + 1 : l0;
    const double lf = lp - l0;
    double wh = fmod(hue + pi, 2.0 * pi);
    if (wh < 0.0) wh += 2.0 * pi;
    const double hp = wh / (2.0 * pi) * 720.0;
    const int h0 = ((int)floor(hp)) % 720, h1 = (h0 + 1) % 720;
    const double hf = hp - floor(hp);
    const double a = cmax[l0 * 720 + h0] + hf * ((double)cmax[l0 * 720 + h1] - cmax[l0 * 720 + h0]);
    const double b = cmax[l1 * 720 + h0] + hf * ((double)cmax[l1 * 720 + h1] - cmax[l1 * 720 + h0]);
    return a + lf * (b - a);
}

static void gamut_compress(const FjOutputColor* oc, double rgb[3]) {
    double xyz[3], lms[3], lab[3];
    mat3(oc->native_rgb_to_d65_xyz, rgb, xyz);
    mat3d(OK_XYZ_LMS, xyz, lms);
    for (int i = 0; i < 3; ++i) lms[i] = cbrt(lms[i]);
    mat3d(OK_LMSR_LAB, lms, lab);
    lab[0] = knee(lab[0], oc->lightness_knee);
    const double chroma = hypot(lab[1], lab[2]);
    if (chroma > 0.0) {
        const double hue = atan2(lab[2], lab[1]);
        const double bnd = fmax(boundary(oc->gamut_cmax.data, lab[0], hue), 1.0e-9);
        const double comp = knee(chroma / bnd, oc->chroma_knee) * bnd;
        lab[1] *= comp / chroma;
        lab[2] *= comp / chroma;
    }
    mat3d(OK_LAB_LMSR, lab, lms);
    for (int i = 0; i < 3; ++i) lms[i] = lms[i] * lms[i] * lms[i];
    mat3d(OK_LMS_XYZ, lms, xyz);
    mat3(oc->d65_xyz_to_native_rgb, xyz, rgb);
}

/* Scanner: normalize -> PCHIP LUT -> exp2 -> correction -> CAT02 -> RGB ->
 * gamut (JuicerCudaScanPipeline.cu:1735-1811). Returns 0 on scan error. */
static int px_scan(const FjoPlan* pl, const double D[3], double glare_percent, double rgb[3]) {
    const FjDensityBounds* sb = &pl->p.scanner_bounds;
    double nrm[3], l2[3];
    for (int c = 0; c < 3; ++c) nrm[c] = (D[c] - (double)sb->min_cmy[c]) * (double)sb->inv_span_cmy[c];
    if (!px_scan_lut(pl, nrm, l2)) return 0;
    double xyz[3] = {exp2(l2[0]), exp2(l2[1]), exp2(l2[2])};
    const FjScannerCorrection* cr = &pl->p.scanner_correction;
    if (cr->flags & FJ_CORRECTION_ACTIVE) {
        const double cy = fmin(fmax((double)cr->xyz_slope * xyz[1] + (double)cr->xyz_offset, 0.0), 1.0);
        const double sc = cy / (xyz[1] + 1e-10);
        for (int c = 0; c < 3; ++c) xyz[c] *= sc;
    }
    /* scanner glare: a veiling fraction of the scan illuminant
     * (JuicerCudaScanPipeline.cu:1801-1806) */
    if (glare_percent != 0.0)
        for (int c = 0; c < 3; ++c) xyz[c] += glare_percent * 0.01 * (double)pl->p.output_color.illuminant_xyz[c];
    double ad[3];
    mat3(pl->p.output_color.cat02, xyz, ad);
    mat3(pl->p.output_color.xyz_to_rgb, ad, rgb);
    if (!isfinite(rgb[0]) || !isfinite(rgb[1]) || !isfinite(rgb[2])) return 0;
    if (pl->p.output_color.flags & FJ_COLOR_GAMUT_COMPRESSION) {
        gamut_compress(&pl->p.output_color, rgb);
        if (!isfinite(rgb[0]) || !isfinite(rgb[1]) || !isfinite(rgb[2])) return 0;
    }
    return 1;
}

/* Output encoding (JuicerCudaScanPipeline.cu:99-190). */
static double encode_channel(const FjOutputEncoding* e, double v) {
    switch (e->cctf_kind) {
        case FJ_CCTF_LINEAR: return v;
        case FJ_CCTF_GAMMA: return copysign(pow(fabs(v), (double)e->gamma), v);
        case FJ_CCTF_SRGB: return v <= 0.0031308 ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
        case FJ_CCTF_BT2020: return v < (double)e->b ? v * 4.5 : (double)e->a * pow(v, 0.45) - ((double)e->a - 1.0);
        case FJ_CCTF_PROPHOTO: return v < (double)e->linear_cutoff ? v * 16.0 : pow(v, (double)e->gamma);
        case FJ_CCTF_DWG_INTERMEDIATE: {
            const double l = fmax(0.0, v);
            return l <= (double)e->linear_cutoff ? l * (double)e->d : (log2(l + (double)e->a) + (double)e->b) * (double)e->c;
        }
        default: return fmin(1.0, fmax(0.0, v));
    }
}

static void px_encode(const FjOutputColor* oc, double rgb[3]) {
    double lin[3];
    if (oc->flags & FJ_COLOR_INPUT_IS_OUTPUT_SPACE) {
        memcpy(lin, rgb, sizeof(lin));
    } else {
        mat3(oc->encoding.dwg_to_output, rgb, lin);
    }
    for (int c = 0; c < 3; ++c) {
        double v = (oc->flags & FJ_COLOR_ENCODE_CCTF) ? encode_channel(&oc->encoding, lin[c]) : lin[c];
        rgb[c] = v <= 0.0 ? 0.0 : (v >= 1.0 ? 1.0 : v);
        if (isnan(v)) rgb[c] = v; /* clamp01d keeps NaN out: NaN compares false */
    }
}

/* ------------------------------------------------------------------------ */
/* spatial operators                                                        */
/* ------------------------------------------------------------------------ */

/* Gaussian FIR weights (rust optics/gaussian.rs:20-56): double samples
 * rounded to float, normalized by the unrounded double sum. */
int fjo_gaussian_weights(int radius, float sigma, float* out) {
    if (radius < 1 || !(sigma > 0.0f) || !isfinite(sigma)) return 1;
    const double s2 = (double)sigma * (double)sigma * 2.0;
    double sum = 0.0;
    for (int i = -radius; i <= radius; ++i) {
        const double v = exp(-((double)i * i) / s2);
        out[i + radius] = (float)v;
        sum += v;
    }
    const double inv = sum != 0.0 ? 1.0 / sum : 0.0;
    for (int i = 0; i <= 2 * radius; ++i) out[i] = (float)((double)out[i] * inv);
    return 0;
}

/* Half-sample reflection with period 2n (JuicerCudaFilmPipeline.cu:82-92). */
static int reflect_periodic(long i, int n) {
    if (n <= 1) return 0;
    const long period = 2L * n;
    long w = i % period;
    if (w < 0) w += period;
    return (int)(w < n ? w : period - 1 - w);
}

/* FIR on a strided line with reflected borders (JuicerCudaFilmPipeline.cu:388-436). */
static void fir_line(const double* in, double* out, int n, ptrdiff_t stride, const float* w, int r) {
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int o = -r; o <= r; ++o) acc += in[reflect_periodic(i + o, n) * stride] * (double)w[o + r];
        out[i * stride] = isfinite(acc) ? acc : 0.0;
    }
}

static void invert3l(long double m[3][3], int size, long double inv[3][3]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inv[r][c] = r == c ? 1.0L : 0.0L;
    for (int col = 0; col < size; ++col) {
        int piv = col;
        for (int r = col + 1; r < size; ++r)
            if (fabsl(m[r][col]) > fabsl(m[piv][col])) piv = r;
        if (piv != col)
            for (int c = 0; c < 3; ++c) {
                long double t = m[piv][c]; m[piv][c] = m[col][c]; m[col][c] = t;
                t = inv[piv][c]; inv[piv][c] = inv[col][c]; inv[col][c] = t;
            }
        const long double sc = m[col][col];
        for (int c = 0; c < size; ++c) {
            m[col][c] /= sc;
            inv[col][c] /= sc;
        }
        for (int r = 0; r < size; ++r) {
            if (r == col) continue;
            const long double f = m[r][col];
            for (int c = 0; c < size; ++c) {
                m[r][c] -= f * m[col][c];
                inv[r][c] -= f * inv[col][c];
            }
        }
    }
}

int fjo_yvv_boundary(const double c[9], int n, double* weights, double* terminal) {
    /* Extended precision (x87 long double) stands in for the reference's
     * double-double arithmetic; the terminal solve is ill-conditioned for
     * long kernels on short lines. */
    const int period = 2 * n;
    long double T[3][3] = {{c[1], c[2], c[3]}, {1, 0, 0}, {0, 1, 0}};
    long double* pw = (long double*)malloc(sizeof(long double) * 3 * (size_t)period);
    pw[0] = c[0];
    pw[1] = 0;
    pw[2] = 0;
    for (int e = 1; e < period; ++e)
        for (int r = 0; r < 3; ++r)
            pw[e * 3 + r] = T[r][0] * pw[(e - 1) * 3] + T[r][1] * pw[(e - 1) * 3 + 1] + T[r][2] * pw[(e - 1) * 3 + 2];
    long double P[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int e = 0; e < period; ++e) {
        long double Q[3][3];
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) Q[r][k] = T[r][0] * P[0][k] + T[r][1] * P[1][k] + T[r][2] * P[2][k];
        memcpy(P, Q, sizeof(P));
    }
    long double S[3][3], Si[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) S[r][k] = (r == k ? 1.0L : 0.0L) - P[r][k];
    invert3l(S, 3, Si);
    for (int s = 0; s < n; ++s) {
        long double rhs[3];
        for (int r = 0; r < 3; ++r) rhs[r] = pw[s * 3 + r] + pw[(period - 1 - s) * 3 + r];
        for (int r = 0; r < 3; ++r) weights[r * n + s] = (double)(Si[r][0] * rhs[0] + Si[r][1] * rhs[1] + Si[r][2] * rhs[2]);
    }
    free(pw);
    const int size = n < 3 ? n : 3;
    long double Ts[3][3] = {{0}}, Ti[3][3];
    for (int r = 0; r < size; ++r) Ts[r][r] = c[4];
    for (int r = 0; r < size; ++r) {
        const long o = n - size + r;
        for (int k = 0; k < 3; ++k) {
            const int col = reflect_periodic(o + k + 1, n) - (n - size);
            if (col >= 0 && col < size) Ts[r][col] -= c[6 + k];
        }
    }
    invert3l(Ts, size, Ti);
    for (int r = 0; r < size; ++r)
        for (int k = 0; k < size; ++k) terminal[r * size + k] = (double)((long double)c[5] * Ti[r][k]);
    return size;
}

/* Reflected YVV recursion on one strided line (JuicerCudaFilmPipeline.cu:478-553
 * operator): the causal then anti-causal recursion applied to the 2n-periodic
 * half-sample reflection of the line. The infinite extension is realized by
 * running both recursions over M reflected samples on either side, M chosen
 * so that the slowest pole has decayed below 1e-18 (equivalent to the
 * reference's solved boundary states without its ill-conditioned solve). */
void fjo_yvv_reflect_line(const double* in, double* out, int n, ptrdiff_t stride, const double c[9]) {
    /* decay length of the recursion's impulse response */
    double s0 = 1.0, s1 = 0.0, s2 = 0.0, peak = 1.0;
    long M = 3;
    for (long m = 0; m < 2000000; ++m) {
        const double v = c[1] * s0 + c[2] * s1 + c[3] * s2;
        s2 = s1;
        s1 = s0;
        s0 = v;
        const double mag = fabs(s0) + fabs(s1) + fabs(s2);
        if (mag > peak) peak = mag;
        if (mag < 1e-18 * peak) {
            M = m + 4;
            break;
        }
        M = m + 4;
    }
    const long total = 2 * M + n;
    double* ext = (double*)malloc(sizeof(double) * (size_t)total);
    double* y = (double*)malloc(sizeof(double) * (size_t)total);
    for (long i = 0; i < total; ++i) ext[i] = in[reflect_periodic(i - M, n) * stride];
    double a = 0, b = 0, d = 0;
    for (long i = 0; i < total; ++i) {
        const double v = c[0] * ext[i] + c[1] * a + c[2] * b + c[3] * d;
        y[i] = v;
        d = b;
        b = a;
        a = v;
    }
    a = b = d = 0;
    for (long i = total - 1; i >= 0; --i) {
        const double v = c[0] * y[i] + c[1] * a + c[2] * b + c[3] * d;
        if (i >= M && i < M + n) out[(i - M) * stride] = v;
        d = b;
        b = a;
        a = v;
    }
    free(ext);
    free(y);
}

/* Weighted sum of separable components on `planes` W x H double planes, in
 * place (JuicerCudaFilmPipeline.cu:798-1478 schedule, FP64 throughout). */
int fjo_dir_filter(const FjSpatialDir* sd, double* const* planes, int nplanes, int W, int H) {
    if (!sd || sd->component_count < 1 || sd->component_count > 4) return 1;
    const size_t N = (size_t)W * H;
    double* acc = (double*)calloc(N, sizeof(double));
    double* tmp = (double*)malloc(sizeof(double) * N);
    double* tmp2 = (double*)malloc(sizeof(double) * N);
    float w[8193];
    for (int p = 0; p < nplanes; ++p) {
        memset(acc, 0, sizeof(double) * N);
        for (size_t ci = 0; ci < sd->component_count; ++ci) {
            const FjDirGaussian* g = &sd->components[ci];
            if (!(g->weight > 0.0f)) continue;
            if (g->reference_operator == FJ_DIR_OPERATOR_IDENTITY) {
                for (size_t i = 0; i < N; ++i) acc[i] += (double)g->weight * planes[p][i];
                continue;
            }
            if (g->reference_operator == FJ_DIR_OPERATOR_FIR_REFLECT) {
                if (g->radius > 4096 || fjo_gaussian_weights(g->radius, g->sigma_px, w)) return 2;
                for (int y = 0; y < H; ++y) fir_line(planes[p] + (size_t)y * W, tmp + (size_t)y * W, W, 1, w, g->radius);
                for (int x = 0; x < W; ++x) fir_line(tmp + x, tmp2 + x, H, W, w, g->radius);
            } else if (g->reference_operator == FJ_DIR_OPERATOR_YVV_REFLECT) {
                const double c[9] = {g->feedforward, g->feedback[0], g->feedback[1], g->feedback[2],
                                     g->normalization_denominator, g->feedforward_numerator,
                                     g->feedback_numerators[0], g->feedback_numerators[1], g->feedback_numerators[2]};
                for (int y = 0; y < H; ++y) fjo_yvv_reflect_line(planes[p] + (size_t)y * W, tmp + (size_t)y * W, W, 1, c);
                for (int x = 0; x < W; ++x) fjo_yvv_reflect_line(tmp + x, tmp2 + x, H, W, c);
            } else {
                return 3;
            }
            for (size_t i = 0; i < N; ++i) acc[i] += (double)g->weight * tmp2[i];
        }
        memcpy(planes[p], acc, sizeof(double) * N);
    }
    free(acc);
    free(tmp);
    free(tmp2);
    return 0;
}

/* Unsharp mask on one W x H double plane (JuicerCudaScanPipeline.cu:316-452,
 * :586-660): half-sample reflected horizontal and vertical blur, non-finite
 * blurs replaced by zero, v + a (v - blur). */
/* ---- scatter / halation (JuicerCudaScatterHalation.cu) ------------------ */

/* One step of the replicate-initialized third-order recursion
 * (JuicerCudaScatterHalation.cu:183-198): legacy coefficients {B, B1, B2, B3}
 * run the direct form, factored {h, g, t, 0} the real pole + second-order
 * section. */
static double yvv_step(int factored, const double c[4], double x, double s[3]) {
    if (factored) {
        s[0] += c[0] * (x - s[0]);
        s[2] = c[2] * s[2] + c[1] * (s[0] - s[1]);
        s[1] += s[2];
        return s[1];
    }
    const double v = c[0] * x + c[1] * s[0] + c[2] * s[1] + c[3] * s[2];
    s[2] = s[1];
    s[1] = s[0];
    s[0] = v;
    return v;
}

static void yvv_init(int factored, double first, double s[3]) {
    s[0] = first;
    s[1] = first;
    s[2] = factored ? 0.0 : first;
}

/* One halation Gaussian on a W x H plane (JuicerCudaScatterHalation.cu:
 * 104-180 FIR vertical-then-horizontal with half-sample reflection,
 * :200-420 YVV horizontal forward/backward then vertical bottom-up/top-down,
 * each pass seeded with its first sample). */
static void hal_gauss(const FjHalationGaussian* g, const double* src, double* dst, double* tmp, int W, int H) {
    const size_t N = (size_t)W * H;
    if (g->kind == FJ_HALATION_IDENTITY) {
        memcpy(dst, src, sizeof(double) * N);
        return;
    }
    if (g->kind == FJ_HALATION_FIR_REFLECT) {
        const int r = g->radius;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                double sum = 0.0;
                for (int o = -r; o <= r; ++o) sum += src[(size_t)reflect_periodic(y - o, H) * W + x] * g->fir_weights[o + r];
                tmp[(size_t)y * W + x] = sum;
            }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                double sum = 0.0;
                for (int o = -r; o <= r; ++o) sum += tmp[(size_t)y * W + reflect_periodic(x + o, W)] * g->fir_weights[o + r];
                dst[(size_t)y * W + x] = sum;
            }
        return;
    }
    const int fac = g->kind == FJ_HALATION_YVV_REPLICATE_FACTORED;
    const double c[4] = {g->yvv[0], g->yvv[1], g->yvv[2], g->yvv[3]};
    double s[3];
    for (int y = 0; y < H; ++y) {
        const double* in = src + (size_t)y * W;
        double* t = tmp + (size_t)y * W;
        yvv_init(fac, in[0], s);
        for (int x = 0; x < W; ++x) t[x] = yvv_step(fac, c, in[x], s);
        yvv_init(fac, t[W - 1], s);
        for (int x = W - 1; x >= 0; --x) t[x] = yvv_step(fac, c, t[x], s);
    }
    for (int x = 0; x < W; ++x) {
        yvv_init(fac, tmp[(size_t)(H - 1) * W + x], s);
        for (int y = H - 1; y >= 0; --y) tmp[(size_t)y * W + x] = yvv_step(fac, c, tmp[(size_t)y * W + x], s);
        yvv_init(fac, tmp[x], s);
        for (int y = 0; y < H; ++y) dst[(size_t)y * W + x] = yvv_step(fac, c, tmp[(size_t)y * W + x], s);
    }
}

/* ---- camera / enlarger diffusion (JuicerCudaDiffusion.cu) ---------------- */

static double ordered_maxd(double v, double floor) { return v < floor ? floor : v; }

/* Nine exponential PSF components of a stage sample
 * (Components::new, rust optics/diffusion_frame.rs:286-420). */
int fjo_diffusion_components(const FjDiffusion* d, double lambda_px[9], double weights[3][9]) {
    static const double kBloomAlpha[4] = {3.2, 3.5, 2.9, 2.5};
    if (d->family > 3 || !(d->pixel_size_um > 0.0) || !(d->spatial_scale > 0.0)) return 2;
    const double centers[3] = {d->group_lambda_um[0], d->group_lambda_um[1], d->group_lambda_um[2]};
    const double spreads[3] = {1.5, 2.0, 2.5};
    const int counts[3] = {2, 3, 4};
    double lam[3][4], w[3][4];
    for (int g = 0; g < 3; ++g) {
        const double lo = log(centers[g] / spreads[g]), hi = log(centers[g] * spreads[g]);
        double total = 0.0;
        for (int i = 0; i < counts[g]; ++i) {
            const double ll = i + 1 == counts[g] ? hi : (i ? lo + (hi - lo) * ((double)i / (counts[g] - 1)) : lo);
            lam[g][i] = exp(ll);
            w[g][i] = g == 2 ? pow(lam[g][i], 2.0 - kBloomAlpha[d->family]) : 1.0;
            total += w[g][i];
        }
        for (int i = 0; i < counts[g]; ++i) w[g][i] /= total;
    }
    /* halo warmth: per-channel reweighting along a centred gradient (:340-373) */
    double grad[3], tt = 0.0, wg = 0.0, halo[3][3];
    for (int i = 0; i < 3; ++i) {
        grad[i] = -1.0 + 2.0 * i / 2.0;
        tt += w[1][i];
        wg += w[1][i] * grad[i];
    }
    wg /= tt;
    for (int i = 0; i < 3; ++i) grad[i] -= wg;
    const double axis[3] = {1.30, 0.15, -1.45};
    for (int ch = 0; ch < 3; ++ch) {
        double tot = 0.0;
        for (int i = 0; i < 3; ++i) {
            halo[ch][i] = ordered_maxd(w[1][i] * (1.0 + d->warmth * axis[ch] * grad[i]), 0.0);
            tot += halo[ch][i];
        }
        for (int i = 0; i < 3; ++i) halo[ch][i] = tot > 0.0 ? halo[ch][i] * (tt / tot) : w[1][i];
    }
    const double scale = ordered_maxd(d->spatial_scale, 1e-6);
    int k = 0;
    for (int g = 0; g < 3; ++g)
        for (int i = 0; i < counts[g]; ++i, ++k) {
            lambda_px[k] = ordered_maxd(lam[g][i] * scale / d->pixel_size_um, 1e-6);
            for (int ch = 0; ch < 3; ++ch)
                weights[ch][k] = g == 1 ? d->group_weights[1] * halo[ch][i] : d->group_weights[g] * w[g][i];
        }
    return 0;
}

/* Normalized PSF samples of one channel on the (2r+1)^2 grid, row-major
 * [dy + r][dx + r]: FP64 sum of nine exponentials stored as f32, divided by
 * the FP64 sum of the stored samples (JuicerCudaDiffusion.cu:294-400). */
int fjo_diffusion_psf(const FjDiffusion* d, int channel, float* psf) {
    double lam[9], wts[3][9];
    if (fjo_diffusion_components(d, lam, wts)) return 2;
    const int r = d->radius_px, n = 2 * r + 1;
    double sum = 0.0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            const double rad = sqrt((double)dx * dx + (double)dy * dy);
            double v = 0.0;
            for (int k = 0; k < 9; ++k) v += wts[channel][k] * exp(-rad / lam[k]) / (2.0 * 3.141592653589793238462643383279502884 * lam[k] * lam[k]);
            const float stored = (float)v;
            psf[(dy + r) * n + dx + r] = stored;
            sum += stored;
        }
    for (int i = 0; i < n * n; ++i) psf[i] = isfinite(sum) && sum > 0.0 ? (float)((double)psf[i] / sum) : NAN;
    return 0;
}

/* Whole-sample mirror (JuicerCudaDiffusion.cu:403-411). */
static int reflect101(long i, int n) {
    const long period = 2L * (n - 1);
    long v = i % period;
    if (v < 0) v += period;
    return (int)(v < n ? v : period - v);
}

int fjo_diffusion(const FjDiffusion* d, double* const planes[3], int W, int H) {
    if (!d || W < 2 || H < 2 || d->radius_px <= 0) return 2;
    const int r = d->radius_px, n = 2 * r + 1;
    const size_t N = (size_t)W * H;
    float* psf = (float*)malloc(sizeof(float) * (size_t)n * n);
    double* conv = (double*)malloc(sizeof(double) * N);
    int* rx = (int*)malloc(sizeof(int) * (size_t)(W + 2 * r));
    int* ry = (int*)malloc(sizeof(int) * (size_t)(H + 2 * r));
    if (!psf || !conv || !rx || !ry) {
        free(psf);
        free(conv);
        free(rx);
        free(ry);
        return 2;
    }
    for (int i = 0; i < W + 2 * r; ++i) rx[i] = reflect101(i - r, W);
    for (int i = 0; i < H + 2 * r; ++i) ry[i] = reflect101(i - r, H);
    const double f = d->scatter_fraction;
    for (int ch = 0; ch < 3; ++ch) {
        if (fjo_diffusion_psf(d, ch, psf)) continue;
        const double* x = planes[ch];
        for (int y = 0; y < H; ++y)
            for (int xx = 0; xx < W; ++xx) {
                double acc = 0.0;
                for (int dy = 0; dy < n; ++dy) {
                    const double* row = x + (size_t)ry[y + dy] * W;
                    const float* pk = psf + (size_t)dy * n;
                    for (int dx = 0; dx < n; ++dx) acc += (double)pk[dx] * row[rx[xx + dx]];
                }
                conv[(size_t)y * W + xx] = acc;
            }
        double* out = planes[ch];
        for (size_t i = 0; i < N; ++i) out[i] = (1.0 - f) * out[i] + f * conv[i];
    }
    free(psf);
    free(conv);
    free(rx);
    free(ry);
    return 0;
}

void fjo_plan_set_fit(FjoPlan* plan, double print_gamma, const float* c, const float* m, const float* y) {
    plan->print_gamma = print_gamma;
    plan->fit_delta[0] = c;
    plan->fit_delta[1] = c ? m : NULL;
    plan->fit_delta[2] = c ? y : NULL;
}

void fjo_plan_capture_grain_input(FjoPlan* plan, float* c, float* m, float* y) {
    plan->grain_capture[0] = c;
    plan->grain_capture[1] = c ? m : NULL;
    plan->grain_capture[2] = c ? y : NULL;
}

void fjo_plan_set_diffusion(FjoPlan* plan, FjoDiffusionFn fn, void* user) {
    plan->diffusion = fn;
    plan->diffusion_user = user;
}

static int plan_diffusion(const FjoPlan* plan, const FjDiffusion* d, double* const planes[3], int W, int H) {
    return plan->diffusion ? plan->diffusion(d, planes, W, H, plan->diffusion_user) : fjo_diffusion(d, planes, W, H);
}

int fjo_halation(const FjScatterHalation* d, double* const planes[3], int W, int H) {
    static const double kTailAmplitude[3] = {0.1633f, 0.6496f, 0.1870f};
    if (!d || W <= 0 || H <= 0) return 2;
    const size_t N = (size_t)W * H;
    double* acc = (double*)malloc(sizeof(double) * N);
    double* g = (double*)malloc(sizeof(double) * N);
    double* tmp = (double*)malloc(sizeof(double) * N);
    if (!acc || !g || !tmp) {
        free(acc);
        free(g);
        free(tmp);
        return 2;
    }
    /* launch_scatter_halation (JuicerCudaScatterHalation.cu:690-852). */
    for (int ch = 0; ch < 3; ++ch) {
        double* x = planes[ch];
        const FjHalationChannel* sc = &d->channels[ch];
        if (d->scatter_amount != 0.0f) {
            const double a = d->scatter_amount, tw = sc->tail_weight;
            const int tail = sc->tail_weight > 0.0f;
            if (tail)
                for (int i = 0; i < 3; ++i) {
                    hal_gauss(&sc->tail[i], x, g, tmp, W, H);
                    for (size_t k = 0; k < N; ++k) acc[k] = (i ? acc[k] : 0.0) + kTailAmplitude[i] * g[k];
                }
            hal_gauss(&sc->core, x, g, tmp, W, H);
            for (size_t k = 0; k < N; ++k) {
                const double scattered = (1.0 - tw) * g[k] + tw * (tail ? acc[k] : 0.0);
                x[k] = (1.0 - a) * x[k] + a * scattered;
            }
        }
        if (sc->total_strength == 0.0f) continue;
        for (uint32_t j = 0; j < d->bounce_count; ++j) {
            hal_gauss(&sc->bounce[j], x, g, tmp, W, H);
            for (size_t k = 0; k < N; ++k) acc[k] = (j ? acc[k] : 0.0) + (double)d->bounce_weights[j] * g[k];
        }
        const double s = sc->total_strength;
        for (size_t k = 0; k < N; ++k) {
            const double added = x[k] + s * (d->bounce_count ? acc[k] : 0.0);
            x[k] = d->renormalize ? added / (1.0 + s) : added;
        }
    }
    free(acc);
    free(g);
    free(tmp);
    return 0;
}

static int reflect_half(long i, int n) {
    if (n <= 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2L * n - i - 1;
    return (int)i;
}

/* Lens blur of one plane: the separable half-sample-reflected Gaussian of
 * blur_plane_in_place (JuicerCudaScanPipeline.cu:2686-2731, kernels :317-430). */
static void lens_plane(double* v, int W, int H, const float* w, int r) {
    const size_t N = (size_t)W * H;
    double* h = (double*)malloc(sizeof(double) * N);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double a = 0.0;
            for (int o = -r; o <= r; ++o) a += v[(size_t)y * W + reflect_half(x + o, W)] * (double)w[o + r];
            h[(size_t)y * W + x] = isfinite(a) ? a : 0.0;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double a = 0.0;
            for (int o = -r; o <= r; ++o) a += h[(size_t)reflect_half(y + o, H) * W + x] * (double)w[o + r];
            v[(size_t)y * W + x] = isfinite(a) ? a : 0.0;
        }
    free(h);
}

/* Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC 2011): ten rounds of the 4x32 S-box with the
 * published multipliers and Weyl key increments; the key is bumped before
 * every round but the first. */
void fjo_philox4x32_10(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]) {
    uint32_t k0 = key[0], k1 = key[1];
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* The glare stream's n-th 32-bit draw at (x, y): OpenRAND's Philox with key
 * {seed high, seed low} and counter {x, y, medium, n}, first output word
 * (third_party/openrand/philox.h:64-125, GlareRngDevice at
 * JuicerCudaScanPipeline.cu:214-228). */
static uint32_t glare_draw(uint64_t seed, uint32_t x, uint32_t y, uint32_t medium, uint32_t n) {
    const uint32_t key[2] = {(uint32_t)(seed >> 32), (uint32_t)seed};
    const uint32_t ctr[4] = {x, y, medium, n};
    uint32_t out[4];
    fjo_philox4x32_10(key, ctr, out);
    return out[0];
}

/* OpenRAND's u01 in FP32: (float)x * 2^-32 + 2^-33 (base_state.h:292-300). */
static float u01f(uint32_t x) { return (float)x * 0x1p-32f + 0x1p-33f; }

/* The glare field (optics_glare_generate_kernel, JuicerCudaScanPipeline.cu:
 * 273-314): per pixel a lognormal sample with mean `percent` and standard
 * deviation roughness x percent from one Box-Muller normal of the pixel's
 * stream; the uniforms are the reference's FP32 words, the rest FP64. */
void fjo_glare_uniforms(uint64_t seed, int64_t abs_x, int64_t abs_y, uint32_t medium, float u[2]) {
    u[0] = u01f(glare_draw(seed, (uint32_t)(uint64_t)abs_x, (uint32_t)(uint64_t)abs_y, medium, 0));
    u[1] = u01f(glare_draw(seed, (uint32_t)(uint64_t)abs_x, (uint32_t)(uint64_t)abs_y, medium, 1));
}

double fjo_glare_sample(uint64_t seed, int64_t abs_x, int64_t abs_y, uint32_t medium, float percent, float roughness) {
    float u[2];
    fjo_glare_uniforms(seed, abs_x, abs_y, medium, u);
    const float u1 = fminf(fmaxf(u[0], 1e-7f), 1.0f), u2 = u[1];
    const double n = sqrt(-2.0 * log((double)u1)) * cos(6.283185307179586476925286766559 * (double)u2);
    const double mean = fmax(0.0, (double)percent), sd = fmax(0.0, (double)roughness * (double)percent);
    const double sigma_sq = log(1.0 + (sd * sd) / (mean * mean));
    const double sigma = sqrt(fmax(0.0, sigma_sq));
    const double mu = log(fmax(1e-12, mean)) - 0.5 * sigma_sq;
    const double g = exp(mu + sigma * n);
    return isfinite(g) ? g : 0.0;
}

/* FNV-1a over the little-endian bytes of 64-bit values
 * (Hash::hash_uint64_values, src/Hash.h:50-56). */
static uint64_t fnv_u64s(const uint64_t* v, int n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < 8; ++b) {
            h ^= (v[i] >> (8 * b)) & 0xFFu;
            h *= 0x100000001b3ull;
        }
    return h;
}

/* The frame's glare seed: recipe hash, frame token and source origin
 * (JuicerCudaExecutor.cpp:2416-2420). */
uint64_t fjo_glare_seed(uint64_t recipe_hash, uint64_t frame_token, int origin_x, int origin_y) {
    const uint64_t v[4] = {recipe_hash, frame_token, (uint64_t)(int64_t)origin_x, (uint64_t)(int64_t)origin_y};
    return fnv_u64s(v, 4);
}

/* Glare is generated when the post descriptor is bound and its percent is
 * finite and positive with a finite roughness (JuicerCudaScanPipeline.cu:
 * 2361-2364, JuicerCudaExecutor.cpp:2405-2413); it is blurred when the glare
 * flag prepared a kernel (ProcessRoot.cpp:3147-3154). */
static int glare_active(const FjScannerPostEffects* sp) {
    return sp->hash != 0 && isfinite(sp->glare_percent) && sp->glare_percent > 0.0f && isfinite(sp->glare_roughness);
}

static int glare_blurred(const FjScannerPostEffects* sp) {
    return (sp->flags & FJ_POST_GLARE) && sp->glare_blur_sigma_px > 0.0f;
}

void fjo_glare_plane(const FjScannerPostEffects* sp, uint64_t seed, uint32_t medium, int origin_x, int origin_y, int W,
                     int H, double* out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            out[(size_t)y * W + x] = fjo_glare_sample(seed, (int64_t)origin_x + x, (int64_t)origin_y + y, medium,
                                                      sp->glare_percent, sp->glare_roughness);
}

/* Gate weave (JuicerCudaScanPipeline.cu:1514-1558, :2122-2146): output pixel
 * (x, y) samples the interleaved RGB planes at R (x - c) + c + (dx, dy), c the
 * window centre ((W-1)/2, (H-1)/2); the position is clamped to the plane and
 * the 4 x 4 Mitchell taps clamp their indices. */
static void weave_planes(const FjFilmEffects* fx, const double* in, double* out, int W, int H) {
    const double cx = 0.5 * (W - 1), cy = 0.5 * (H - 1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double ox = x - cx, oy = y - cy;
            double sx = (double)fx->weave_cos_rot * ox - (double)fx->weave_sin_rot * oy + cx + fx->weave_dx_px;
            double sy = (double)fx->weave_sin_rot * ox + (double)fx->weave_cos_rot * oy + cy + fx->weave_dy_px;
            sx = isnan(sx) ? 0.0 : fmin(fmax(sx, 0.0), W - 1.0);
            sy = isnan(sy) ? 0.0 : fmin(fmax(sy, 0.0), H - 1.0);
            const int ix = (int)floor(sx), iy = (int)floor(sy);
            double wx[4], wy[4];
            for (int k = 0; k < 4; ++k) {
                wx[k] = mitchell(sx - ix + 1.0 - k);
                wy[k] = mitchell(sy - iy + 1.0 - k);
            }
            double acc[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < 4; ++j) {
                int yy = iy + j - 1;
                yy = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
                for (int i = 0; i < 4; ++i) {
                    int xx = ix + i - 1;
                    xx = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
                    for (int c = 0; c < 3; ++c) acc[c] += wy[j] * wx[i] * in[((size_t)yy * W + xx) * 3 + c];
                }
            }
            for (int c = 0; c < 3; ++c) out[((size_t)y * W + x) * 3 + c] = acc[c];
        }
}

static void unsharp_plane(double* v, int W, int H, const float* w, int r, double amount) {
    const size_t N = (size_t)W * H;
    double* h = (double*)malloc(sizeof(double) * N);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double a = 0.0;
            for (int o = -r; o <= r; ++o) a += v[(size_t)y * W + reflect_half(x + o, W)] * (double)w[o + r];
            h[(size_t)y * W + x] = isfinite(a) ? a : 0.0;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double a = 0.0;
            for (int o = -r; o <= r; ++o) a += h[(size_t)reflect_half(y + o, H) * W + x] * (double)w[o + r];
            const double b = isfinite(a) ? a : 0.0;
            const double v0 = v[(size_t)y * W + x];
            const double o = v0 + amount * (v0 - b);
            v[(size_t)y * W + x] = isfinite(o) ? o : 0.0;
        }
    free(h);
}

/* ------------------------------------------------------------------------ */
/* schedule                                                                 */
/* ------------------------------------------------------------------------ */

typedef void (*RowFn)(Frame* f, int y0, int y1);

typedef struct Job {
    Frame* f;
    RowFn fn;
    int y0, y1;
} Job;

static void* job_main(void* arg) {
    Job* j = (Job*)arg;
    j->fn(j->f, j->y0, j->y1);
    return NULL;
}

static void run_rows(Frame* f, RowFn fn, int threads) {
    if (threads <= 1 || f->h < 2 * threads) {
        fn(f, 0, f->h);
        return;
    }
    pthread_t tid[256];
    Job jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; ++t) {
        jobs[t].f = f;
        jobs[t].fn = fn;
        jobs[t].y0 = (int)((long)f->h * t / threads);
        jobs[t].y1 = (int)((long)f->h * (t + 1) / threads);
        pthread_create(&tid[t], NULL, job_main, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

static void add_errors(Frame* f, uint32_t bits) {
    if (!bits) return;
    pthread_mutex_lock(&f->lock);
    f->errors |= bits;
    pthread_mutex_unlock(&f->lock);
}

/* Print + scan + colour of one developed pixel into the linear planes. */
static uint32_t px_back(const FjoPlan* pl, Frame* f, size_t idx, double D[3]) {
    double rgb[3];
    if (f->dens_out) {
        /* visual grain follows on the density planes (develop_print_capture,
         * JuicerCudaExecutor.cpp:2198-2280) */
        for (int c = 0; c < 3; ++c) {
            if (f->dens64[0]) f->dens64[c][idx] = D[c];
            else f->dens[c][idx] = (float)D[c];
        }
        return 0;
    }
    if (pl->print && f->plin[0]) {
        /* enlarger diffusion follows: keep the print-linear exposure */
        double lin[3];
        px_print_linear(pl, D, lin);
        for (int c = 0; c < 3; ++c) f->plin[c][idx] = lin[c];
        return 0;
    }
    if (pl->print) px_print(pl, D);
    const int ok = px_scan(pl, D, f->glare ? f->glare[idx] : 0.0, rgb);
    if (!ok) rgb[0] = rgb[1] = rgb[2] = 0.0;
    memcpy(f->linear + idx * 3, rgb, sizeof(rgb));
    return ok ? 0u : 1u;
}

/* Print development from the diffused enlarger exposure, then the scanner
 * (develop_print_density_from_enlarger_linear_kernel,
 * JuicerCudaScanPipeline.cu:1449-1494; print_log_encode_diffused_device). */
static void rows_enlarged(Frame* f, int y0, int y1) {
    const FjoPlan* pl = f->plan;
    uint32_t errs = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            const size_t idx = (size_t)y * f->w + x;
            double logp[3], D[3], rgb[3];
            for (int c = 0; c < 3; ++c) logp[c] = log10(fmax(f->plin[c][idx], 0.0) + 1e-10);
            print_curves(pl, logp, D);
            const int ok = px_scan(pl, D, f->glare ? f->glare[idx] : 0.0, rgb);
            if (!ok) rgb[0] = rgb[1] = rgb[2] = 0.0;
            memcpy(f->linear + idx * 3, rgb, sizeof(rgb));
            if (!ok) errs |= 1u;
        }
    add_errors(f, errs);
}

/* Back end from the grained density planes. */
static void rows_density(Frame* f, int y0, int y1) {
    uint32_t errs = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            const size_t idx = (size_t)y * f->w + x;
            double D[3];
            for (int c = 0; c < 3; ++c) D[c] = f->dens64[0] ? f->dens64[c][idx] : (double)f->dens[c][idx];
            errs |= px_back(f->plan, f, idx, D);
        }
    add_errors(f, errs);
}

/* Non-spatial chain: exposure -> development -> back end. */
static void rows_chain(Frame* f, int y0, int y1) {
    const FjoPlan* pl = f->plan;
    uint32_t errs = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            double E[3], D[3];
            if (f->carrier[0])
                carrier_exposure(pl, f->carrier, (size_t)y * f->w + x, E);
            else
                px_exposure(pl, f->ae, f->src + y * f->src_stride + (ptrdiff_t)x * f->comps, E);
            errs |= px_develop(pl, E, D);
            errs |= px_back(pl, f, (size_t)y * f->w + x, D);
        }
    add_errors(f, errs);
}

/* Camera carrier planes (camera_film_linear_exposure_kernel,
 * JuicerCudaFilmPipeline.cu:3103-3146). */
static void rows_carrier(Frame* f, int y0, int y1) {
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            double E[3];
            px_camera(f->plan, f->ae, f->src + y * f->src_stride + (ptrdiff_t)x * f->comps, E);
            const size_t idx = (size_t)y * f->w + x;
            for (int c = 0; c < 3; ++c) f->carrier[c][idx] = E[c];
        }
}

/* Spatial-DIR source (JuicerCudaFilmPipeline.cu:124-220): raw donor
 * corrections (Y/M/C) and unsanitized log raw exposure (B/G/R) planes. */
static void rows_dir_source(Frame* f, int y0, int y1) {
    const FjoPlan* pl = f->plan;
    const FjDirCouplers* dir = &pl->p.dir_couplers;
    const FjFilmDevelopment* fd = &pl->p.film_development;
    Curve dens[3], comp[3];
    film_curves(pl, dens, comp);
    const double gamma[3] = {fd->gamma_rgb[2], fd->gamma_rgb[1], fd->gamma_rgb[0]};
    double M[9], dmax[3], dref[3], dk[3];
    for (int d = 0; d < 3; ++d) {
        for (int r = 0; r < 3; ++r) M[d * 3 + r] = dir->matrix_rgb[(2 - d) * 3 + (2 - r)];
        dmax[d] = dir->density_max_rgb[2 - d];
        dref[d] = dir->density_ref_rgb[2 - d];
        dk[d] = dir->donor_k_rgb[2 - d];
    }
    uint32_t errs = 0;
    const size_t N = (size_t)f->w * f->h;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            const size_t idx = (size_t)y * f->w + x;
            double E[3], layer[3], donors[3];
            if (f->carrier[0])
                carrier_exposure(pl, f->carrier, idx, E);
            else
                px_exposure(pl, f->ae, f->src + y * f->src_stride + (ptrdiff_t)x * f->comps, E);
            uint32_t fail = 0;
            for (int b = 0; b < 3; ++b) {
                const double raw = log10(fmax(E[b], 0.0) + 1e-10);
                f->plane[3 + b][idx] = raw;
                layer[b] = curve_at(&dens[b], sanitize_inf_logE(raw, &dens[b]), gamma[b]);
            }
            for (int c = 0; c < 3; ++c) {
                donors[c] = dmax[c] - layer[c];
                if (dir->mode == FJ_DIR_NEGATIVE_DONOR_LANGMUIR && !langmuir(layer[c], dk[c], dref[c], &donors[c])) fail |= 1u << c;
            }
            for (int r = 0; r < 3; ++r) {
                const double v = M[r] * donors[0] + M[3 + r] * donors[1] + M[6 + r] * donors[2];
                if (!isfinite(v)) fail |= 1u << r;
                f->plane[r][idx] = v;
            }
            if (fail) errs |= (1u << 8) | ((fail & 7u) << 9);
        }
    (void)N;
    add_errors(f, errs);
}

/* Final development against filtered corrections
 * (JuicerCudaDeviceHelpers.cuh:1255-1329), then the back end. */
static void rows_dir_final(Frame* f, int y0, int y1) {
    const FjoPlan* pl = f->plan;
    const FjDirCouplers* dir = &pl->p.dir_couplers;
    const FjFilmDevelopment* fd = &pl->p.film_development;
    Curve dens[3], comp[3];
    film_curves(pl, dens, comp);
    const double gamma[3] = {fd->gamma_rgb[2], fd->gamma_rgb[1], fd->gamma_rgb[0]};
    double cref[3], kr[3];
    for (int d = 0; d < 3; ++d) {
        cref[d] = dir->receiver_c_ref_rgb[2 - d];
        kr[d] = dir->receiver_kr_rgb[2 - d];
    }
    uint32_t errs = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < f->w; ++x) {
            const size_t idx = (size_t)y * f->w + x;
            double corr[3], l[3], D[3];
            uint32_t fail = 0;
            for (int c = 0; c < 3; ++c) corr[c] = f->plane[c][idx];
            if (dir->mode == FJ_DIR_POSITIVE_RECEIVER_LANGMUIR)
                for (int c = 0; c < 3; ++c) {
                    if (cref[c] == 0.0) {
                        corr[c] = 0.0;
                        continue;
                    }
                    double v;
                    if (!langmuir(corr[c], kr[c], cref[c], &v)) fail |= 1u << c;
                    corr[c] = v;
                }
            if (!fail)
                for (int c = 0; c < 3; ++c) {
                    const double o = f->plane[3 + c][idx];
                    if (isnan(o)) {
                        fail |= 1u << c;
                        continue;
                    }
                    l[c] = o - corr[c];
                    if (!isfinite(l[c]) && isfinite(o)) fail |= 1u << c;
                }
            if (fail) {
                errs |= (1u << 8) | ((fail & 7u) << 9);
                D[0] = D[1] = D[2] = NAN;
            } else {
                for (int c = 0; c < 3; ++c) l[c] = sanitize_inf_logE(l[c], &dens[c]);
                D[2] = curve_at(&comp[0], l[0], gamma[0]);
                D[1] = curve_at(&comp[1], l[1], gamma[1]);
                D[0] = curve_at(&comp[2], l[2], gamma[2]);
            }
            errs |= px_back(pl, f, idx, D);
        }
    add_errors(f, errs);
}

int fjo_render(const FjoPlan* plan, const float* src, int width, int height, int comps,
               ptrdiff_t src_stride, float* dst, ptrdiff_t dst_stride, double* linear, int threads,
               uint32_t* error_flags) {
    if (!plan || !src || width <= 0 || height <= 0 || (comps != 3 && comps != 4)) return 2;
    const FjPreparedHostData* p = &plan->p;
    const int grain = (p->grain.flags & FJ_GRAIN_ACTIVE) != 0;
    const int halation = (p->optics.flags & FJ_OPTICS_SCATTER_HALATION) != 0;
    const int camera_diffusion = (p->optics.flags & FJ_OPTICS_CAMERA_DIFFUSION) != 0;
    const int enlarger_diffusion = plan->print && (p->optics.flags & FJ_OPTICS_ENLARGER_DIFFUSION) != 0;
    const size_t N = (size_t)width * height;
    Frame f;
    memset(&f, 0, sizeof(f));
    f.plan = plan;
    f.src = src;
    f.w = width;
    f.h = height;
    f.comps = comps;
    f.src_stride = src_stride;
    f.dst = dst;
    f.dst_stride = dst_stride;
    f.ae = fjo_auto_exposure(plan, src, width, height, comps, src_stride);
    if (f.ae < 0.0) return 5; /* auto exposure over a source other than the frame */
    f.linear = (double*)malloc(sizeof(double) * N * 3);
    pthread_mutex_init(&f.lock, NULL);
    int rc = 0;
    const int boost = (p->optics.flags & FJ_OPTICS_HIGHLIGHT_BOOST) != 0;
    const FjScannerPostEffects* sp = &p->scanner_post;
    /* A bound post descriptor prepares each Gaussian whose sigma is positive
     * and refuses a non-finite sigma or a radius outside 1..75
     * (ProcessRoot.cpp:3121-3158, JuicerCudaResources.cpp:6436-6448). */
    if (sp->hash != 0 &&
        ((sp->lens_blur_sigma_px > 0.0f &&
          (!isfinite(sp->lens_blur_sigma_px) || sp->lens_blur_radius < 1 || sp->lens_blur_radius > 75)) ||
         (sp->unsharp_sigma_px > 0.0f &&
          (!isfinite(sp->unsharp_sigma_px) || sp->unsharp_radius < 1 || sp->unsharp_radius > 75)) ||
         (glare_blurred(sp) &&
          (!isfinite(sp->glare_blur_sigma_px) || sp->glare_blur_radius < 1 || sp->glare_blur_radius > 75)))) {
        free(f.linear);
        pthread_mutex_destroy(&f.lock);
        return 4;
    }
    if (glare_active(sp)) {
        /* generated and blurred before the scanner reads it
         * (launch_focused_scan_linear_rgb, JuicerCudaScanPipeline.cu:2361-2405) */
        const uint64_t seed = fjo_glare_seed(p->recipe_hash, plan->frame_token, plan->origin_x, plan->origin_y);
        f.glare = (double*)malloc(sizeof(double) * N);
        fjo_glare_plane(sp, seed, plan->print ? 1u : 0u, plan->origin_x, plan->origin_y, width, height, f.glare);
        if (glare_blurred(sp)) {
            float* w = (float*)malloc(sizeof(float) * (2 * (size_t)sp->glare_blur_radius + 1));
            if (fjo_gaussian_weights(sp->glare_blur_radius, sp->glare_blur_sigma_px, w)) rc = 4;
            else lens_plane(f.glare, width, height, w, sp->glare_blur_radius);
            free(w);
        }
    }
    if (rc == 0 && (halation || camera_diffusion || boost)) {
        /* expose_print_camera (JuicerCudaExecutor.cpp:1874-1989): carrier,
         * highlight boost, camera diffusion, then scatter/halation in place
         * on it. */
        for (int c = 0; c < 3; ++c) f.carrier[c] = (double*)malloc(sizeof(double) * N);
        run_rows(&f, rows_carrier, threads);
        double* const rgb_planes[3] = {f.carrier[2], f.carrier[1], f.carrier[0]};
        if (boost) rc = fjo_highlight_boost(&p->optics.highlight_boost, rgb_planes, N);
        if (rc == 0 && camera_diffusion) rc = plan_diffusion(plan, &p->optics.camera, rgb_planes, width, height);
        if (rc == 0 && halation) rc = fjo_halation(&p->optics.scatter_halation, rgb_planes, width, height);
    }
    if (enlarger_diffusion)
        for (int c = 0; c < 3; ++c) f.plin[c] = (double*)malloc(sizeof(double) * N);
    const int fixed_grain = plan->fit_delta[0] != NULL;
    if (grain || fixed_grain) {
        for (int c = 0; c < 3; ++c) {
            if (fixed_grain) f.dens64[c] = (double*)malloc(sizeof(double) * N);
            else f.dens[c] = (float*)malloc(sizeof(float) * N);
        }
        f.dens_out = 1;
    }
    if (rc != 0) {
    } else if (p->spatial_dir.hash != 0) {
        for (int i = 0; i < 6; ++i) f.plane[i] = (double*)malloc(sizeof(double) * N);
        run_rows(&f, rows_dir_source, threads);
        rc = fjo_dir_filter(&p->spatial_dir, f.plane, 3, width, height);
        if (rc == 0) run_rows(&f, rows_dir_final, threads);
        for (int i = 0; i < 6; ++i) free(f.plane[i]);
    } else {
        run_rows(&f, rows_chain, threads);
    }
    if (rc == 0 && (grain || fixed_grain)) {
        if (plan->grain_capture[0] && !fixed_grain)
            for (int c = 0; c < 3; ++c) memcpy(plan->grain_capture[c], f.dens[c], sizeof(float) * N);
        if (grain && !fixed_grain) rc = fjo_grain_apply(p, f.dens, width, height);
        if (rc == 0 && fixed_grain)
            for (int c = 0; c < 3; ++c)
                for (size_t i = 0; i < N; ++i) f.dens64[c][i] += (double)plan->fit_delta[c][i];
        f.dens_out = 0;
        if (rc == 0) run_rows(&f, rows_density, threads);
    }
    if (rc == 0 && enlarger_diffusion) {
        /* develop_print_medium (JuicerCudaExecutor.cpp:2282-2382). */
        rc = plan_diffusion(plan, &p->optics.enlarger, f.plin, width, height);
        if (rc == 0) run_rows(&f, rows_enlarged, threads);
    }
    if (rc == 0 && sp->hash != 0 && sp->lens_blur_sigma_px > 0.0f) {
        /* lens blur before the unsharp mask (JuicerCudaScanPipeline.cu:2795-2809) */
        float* w = (float*)malloc(sizeof(float) * (2 * (size_t)sp->lens_blur_radius + 1));
        if (fjo_gaussian_weights(sp->lens_blur_radius, sp->lens_blur_sigma_px, w)) {
            rc = 4;
        } else {
            double* pl3 = (double*)malloc(sizeof(double) * N);
            for (int c = 0; c < 3; ++c) {
                for (size_t i = 0; i < N; ++i) pl3[i] = f.linear[i * 3 + c];
                lens_plane(pl3, width, height, w, sp->lens_blur_radius);
                for (size_t i = 0; i < N; ++i) f.linear[i * 3 + c] = pl3[i];
            }
            free(pl3);
        }
        free(w);
    }
    if (rc == 0 && sp->hash != 0 && sp->unsharp_sigma_px > 0.0f && sp->unsharp_radius > 0 &&
        isfinite(sp->unsharp_amount) && sp->unsharp_amount > 0.0f) {
        /* JuicerCudaScanPipeline.cu:2811-2866 */
        float w[151];
        if (fjo_gaussian_weights(sp->unsharp_radius, sp->unsharp_sigma_px, w)) {
            rc = 4;
        } else {
            double* pl3 = (double*)malloc(sizeof(double) * N);
            for (int c = 0; c < 3; ++c) {
                for (size_t i = 0; i < N; ++i) pl3[i] = f.linear[i * 3 + c];
                unsharp_plane(pl3, width, height, w, sp->unsharp_radius, sp->unsharp_amount);
                for (size_t i = 0; i < N; ++i) f.linear[i * 3 + c] = pl3[i];
            }
            free(pl3);
        }
    }
    if (rc == 0 && (p->effects.flags & FJ_EFFECTS_GATE_OUTPUT) && (p->effects.flags & FJ_EFFECTS_WEAVE) &&
        p->effects.hash != 0 && !(grain && p->grain.debug_view != 0)) {
        /* gate weave on the post-unsharp planes (JuicerCudaScanPipeline.cu:2122-2146) */
        double* woven = (double*)malloc(sizeof(double) * N * 3);
        weave_planes(&p->effects, f.linear, woven, width, height);
        free(f.linear);
        f.linear = woven;
    }
    if (rc == 0) {
        if (linear) memcpy(linear, f.linear, sizeof(double) * N * 3);
        if (dst)
            for (int y = 0; y < height; ++y)
                for (int x = 0; x < width; ++x) {
                    double rgb[3];
                    memcpy(rgb, f.linear + ((size_t)y * width + x) * 3, sizeof(rgb));
                    px_encode(&p->output_color, rgb);
                    float* o = dst + y * dst_stride + (ptrdiff_t)x * comps;
                    const float* px = src + y * src_stride + (ptrdiff_t)x * comps;
                    o[0] = (float)rgb[0];
                    o[1] = (float)rgb[1];
                    o[2] = (float)rgb[2];
                    if (comps == 4) o[3] = px[3];
                }
    }
    for (int c = 0; c < 3; ++c) {
        free(f.carrier[c]);
        free(f.plin[c]);
        free(f.dens[c]);
        free(f.dens64[c]);
    }
    free(f.glare);
    free(f.linear);
    pthread_mutex_destroy(&f.lock);
    if (error_flags) *error_flags = f.errors;
    return rc;
}

void fjo_plan_set_frame(FjoPlan* plan, uint64_t frame_token, int origin_x, int origin_y) {
    plan->frame_token = frame_token;
    plan->origin_x = origin_x;
    plan->origin_y = origin_y;
}

/* ------------------------------------------------------------------------ */
/* per-pixel stages as test entry points                                    */
/* ------------------------------------------------------------------------ */

void fjo_film_raw(const FjFilmExposure* fe, const float rgb[3], double xyz[3], double srgb[3], double sums_bgr[3],
                  double raw_bgr[3]) {
    film_raw(fe, rgb, xyz, srgb, sums_bgr, raw_bgr);
}

void fjo_tc_exposure(const float* tc_lut_rgba, const double xyz[3], double raw_bgr[3]) {
    const double w[3] = {fin0(xyz[0]), fin0(xyz[1]), fin0(xyz[2])};
    tc_exposure(tc_lut_rgba, w, raw_bgr);
}

void fjo_print_integral(const FjPrint* print, const double density_cmy[3], double raw[3]) {
    print_integral(print, density_cmy, raw);
}

double fjo_curve_at(const float* x, const float* y, int n, int stride, double log_exposure, double gamma) {
    const Curve cv = {x, y, n, stride};
    return curve_at(&cv, log_exposure, gamma);
}

int fjo_langmuir(double v, double k, double ref, double* out) { return langmuir(v, k, ref, out); }

int fjo_plan_scan_lut(const FjoPlan* plan, const double normalized_cmy[3], double log2_xyz[3]) {
    return px_scan_lut(plan, normalized_cmy, log2_xyz);
}

int fjo_plan_scan(const FjoPlan* plan, const double density_cmy[3], double rgb[3]) {
    return px_scan(plan, density_cmy, 0.0, rgb);
}

void fjo_gamut_compress(const FjOutputColor* color, double rgb[3]) { gamut_compress(color, rgb); }

void fjo_encode(const FjOutputColor* color, double rgb[3]) { px_encode(color, rgb); }

/* FNV-1a 64 of a byte string (the reference's table digests,
 * tests/ffi/tc_lut_test.cpp:71-80). */
uint64_t fjo_fnv1a64(const void* data, size_t bytes) {
    const unsigned char* p = (const unsigned char*)data;
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
