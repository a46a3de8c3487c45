/*
 * fjo_grain.c — CPU restatement of the visual-grain stage (density space).
 *
 * TEST INFRASTRUCTURE ONLY (see fj_oracle.h).
 *
 * Unlike the rest of the oracle this stage is restated in FP32, operation for
 * operation: its particle counts are Poisson samples, discontinuous in their
 * inputs, so an FP64 restatement would disagree with any FP32 implementation
 * on a fraction of the samples. Evaluating the same FP32 arithmetic (no
 * contraction: this file is built with -ffp-contract=off) and the same
 * deterministic exp/log/sin/cos (Cody–Waite reduction + Taylor polynomials,
 * checked against libm by the tests) makes every count — and so every RNG
 * draw — reproducible.
 *
 * Sources restated:
 *   payload          pack_visual_grain_payload (src/Cuda/JuicerCudaFilmPayloads.cpp:212-500)
 *   frame uniforms   grain_prepare_frame_uniforms_kernel (src/Cuda/Film/JuicerCudaFilmPipeline.cu:2174-2352)
 *   RNG / noise      splitmix64, STBN and Wang addressing, value noise (…FilmPipeline.cu:1602-1995)
 *   sampler          fast_poisson_device, layer_particle_model_device,
 *                    interp_density_layer_device (…FilmPipeline.cu:1997-2172)
 *   pixel kernels    grain_apply_simple / grain_layer / grain_layer_triplet (…FilmPipeline.cu:2579-2987)
 *   schedule         juicer_cuda_apply_visual_grain (src/Cuda/Scan/JuicerCudaScanPipeline.cu:3141-3756)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fj_oracle.h"

#define GOLDEN 0x9E3779B97F4A7C15ULL
#define MIX1 0xBF58476D1CE4E5B9ULL
#define MIX2 0x94D049BB133111EBULL

/* ---- deterministic FP32 elementary functions ------------------------------ */
float fjo_det_expf(float x) {
    if (isnan(x)) return x;
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
    return ldexpf(p, (int)k);
}

float fjo_det_logf(float x) {
    if (isnan(x) || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (isinf(x)) return x;
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
    const float ef = (float)e;
    return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.428606765330187045037746429443359375e-06f, lm));
}

static void det_sincos(float x, float* s_out, float* c_out) {
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
    const int q = ((int)k) & 3;
    switch (q) {
        case 0: *s_out = s; *c_out = c; break;
        case 1: *s_out = c; *c_out = -s; break;
        case 2: *s_out = -s; *c_out = -c; break;
        default: *s_out = -c; *c_out = s; break;
    }
}

float fjo_det_cosf(float x) {
    float s, c;
    det_sincos(x, &s, &c);
    return c;
}

float fjo_det_sinf(float x) {
    float s, c;
    det_sincos(x, &s, &c);
    return s;
}

/* ---- integer streams ------------------------------------------------------ */
uint64_t fjo_splitmix64(uint64_t v) {
    v += GOLDEN;
    v = (v ^ (v >> 30)) * MIX1;
    v = (v ^ (v >> 27)) * MIX2;
    return v ^ (v >> 31);
}

static uint64_t fnv_u64s(const uint64_t* v, int n) {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < 8; ++b) {
            h ^= (v[i] >> (8 * b)) & 0xFFu;
            h *= 0x100000001B3ULL;
        }
    return h;
}

static int wrap_coord(int64_t c, int size) {
    if (size <= 0) return 0;
    if ((size & (size - 1)) == 0) return (int)((uint64_t)c & (uint64_t)(size - 1));
    c %= (int64_t)size;
    if (c < 0) c += size;
    return (int)c;
}

static int is_finite_f(float v) { return isfinite(v); }

/* ---- payload ------------------------------------------------------------- */
typedef struct {
    float* w;  /* 2r+1 weights, or NULL */
    int r;
    float sigma;
} Fir;

typedef struct {
    /* noise */
    const uint8_t* stbn;
    int stbn_w, stbn_h, stbn_frames;
    int stbn_off_x, stbn_off_y, stbn_frame;
    int64_t frame_index;
    const uint8_t* wang_tiles;
    const uint8_t* wang_lut;
    int wang_w, wang_h, wang_count, wang_colors;
    float wang_cell_mm, pixel_size_um;
    uint64_t session_seed, clip_token;
    int origin_x, origin_y, pitch_px;
    float time_alpha;
    /* particle model */
    uint64_t seed_base, seed_base_next;
    float density_min[3], density_max[3], n_particles[3], od_particle[3], uniformity[3];
    float dmin_l[3][3], dmax_l[3][3], n_l[3][3], od_l[3][3];
    int n_sublayers, positive_film, sublayers;
    const float* base_curve[3];
    int curve_n;
    const float* layer_curve[3][3];
    int axis_finite[3];
    float axis_prefix[3][16];
    float size_mix_weight, size_mix_scale; /* coarse weight and scale inside a pass */
    /* temporal constants */
    int breathing_period, clump_period;
    float breathing_amp, breathing_cell_small_um, breathing_cell_large_um, breathing_mix, breathing_drift;
    float micro[2], clump_temporal_mix;
    /* frame uniforms */
    int breathing_active, clump_active;
    float breathing_t, breathing_cell_small_px, breathing_cell_large_px, breathing_drift_x, breathing_drift_y,
        breathing_roll_y, breathing_mix_u;
    uint64_t breathing_seed_a, breathing_seed_b, clump_seed;
    float clump_cell_px, clump_roll_y, clump_rms_norm, clump_mu, clump_sigma;
} Grain;

static float smoothstep_f(float t) { return t * t * (3.0f - 2.0f * t); }
static float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

static float hash01(int x, int y, uint64_t seed) {
    const uint64_t h = fjo_splitmix64(seed ^ ((uint64_t)(int64_t)x * GOLDEN) ^ ((uint64_t)(int64_t)y * MIX1));
    return (float)(uint32_t)(h & 0xFFFFFFFFu) * (1.0f / 4294967296.0f);
}

static float spatial_hash01(int x, int y, uint64_t seed) {
    const uint32_t lo = (uint32_t)seed, hi = (uint32_t)(seed >> 32);
    uint32_t h = lo ^ ((hi << 16) | (hi >> 16));
    h ^= (uint32_t)x * 0x9E3779B9u;
    h ^= (uint32_t)y * 0x85EBCA6Bu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return (float)h * (1.0f / 4294967296.0f);
}

static uint32_t hash_u32(int x, int y, uint64_t seed, uint64_t salt) {
    const uint64_t h =
        fjo_splitmix64(seed ^ salt ^ ((uint64_t)(int64_t)x * GOLDEN) ^ ((uint64_t)(int64_t)y * MIX1));
    return (uint32_t)(h & 0xFFFFFFFFu);
}

static float value_noise(float x, float y, uint64_t seed) {
    const int ix = (int)floorf(x), iy = (int)floorf(y);
    const float fx = x - (float)ix, fy = y - (float)iy;
    const float sx = smoothstep_f(fx), sy = smoothstep_f(fy);
    const float v00 = spatial_hash01(ix, iy, seed), v10 = spatial_hash01(ix + 1, iy, seed);
    const float v01 = spatial_hash01(ix, iy + 1, seed), v11 = spatial_hash01(ix + 1, iy + 1, seed);
    const float v0 = v00 + (v10 - v00) * sx;
    const float v1 = v01 + (v11 - v01) * sx;
    return v0 + (v1 - v0) * sy;
}

static float lognormal(float mean, float sd, float n) {
    const float m2 = mean * mean, s2 = sd * sd;
    const float sigma_sq = fjo_det_logf(1.0f + (s2 / m2));
    const float sigma = sqrtf(fmaxf(0.0f, sigma_sq));
    const float mu = fjo_det_logf(fmaxf(1e-12f, mean)) - 0.5f * sigma_sq;
    return fjo_det_expf(mu + sigma * n);
}

static void frame_uniforms(Grain* g) {
    const float time = (float)g->frame_index + g->time_alpha;
    const float two_pi = 6.28318530717958647692f;
    g->breathing_active = 0;
    if (g->breathing_amp > 0.0f && g->breathing_period > 0 && g->pixel_size_um > 0.0f &&
        g->breathing_cell_small_um > 0.0f && g->breathing_cell_large_um > 0.0f) {
        const float period = (float)g->breathing_period;
        const float step_f = floorf(time / period);
        const int64_t step = (int64_t)step_f;
        const float frac = (time - step_f * period) / period;
        g->breathing_t = smoothstep_f(clampf(frac, 0.0f, 1.0f));
        const uint64_t seed = g->session_seed ? g->session_seed : 1ULL;
        g->breathing_seed_a = fjo_splitmix64(seed ^ ((uint64_t)step * 0x8EBC6AF09C88C6E3ULL));
        g->breathing_seed_b = fjo_splitmix64(seed ^ ((uint64_t)(step + 1) * 0x8EBC6AF09C88C6E3ULL));
        const float inv_pixel = 1.0f / g->pixel_size_um;
        g->breathing_cell_small_px = fmaxf(4.0f, g->breathing_cell_small_um * inv_pixel);
        g->breathing_cell_large_px = fmaxf(4.0f, g->breathing_cell_large_um * inv_pixel);
        float drift_px = 0.0f;
        if (g->breathing_drift > 0.0f) drift_px = g->breathing_drift * inv_pixel;
        float dx = 0.0f, dy = 0.0f;
        if (drift_px > 0.0f) {
            const uint64_t h = fjo_splitmix64(seed ^ 0x6A09E667F3BCC909ULL);
            const float angle = (float)(uint32_t)(h & 0xFFFFFFFFu) * (1.0f / 4294967296.0f) * two_pi;
            dx = fjo_det_cosf(angle) * drift_px;
            dy = fjo_det_sinf(angle) * drift_px;
        }
        const float roll = g->pitch_px > 0 ? (float)g->pitch_px : 0.0f;
        g->breathing_drift_x = dx * time;
        g->breathing_drift_y = dy * time;
        g->breathing_roll_y = roll * time;
        g->breathing_mix_u = clampf(g->breathing_mix, 0.0f, 1.0f);
        g->breathing_active = 1;
    }
    g->clump_active = 0;
    const float sd_spatial = g->micro[1] * 0.001f;
    if (sd_spatial > 0.0f && g->pixel_size_um > 0.0f) {
        float cell_um = g->micro[0];
        if (!(cell_um > 0.0f)) cell_um = 0.0f;
        const float inv_pixel = 1.0f / g->pixel_size_um;
        g->clump_cell_px = cell_um * inv_pixel;
        if (!(g->clump_cell_px > 4.0f)) g->clump_cell_px = 4.0f;
        const float roll = g->pitch_px > 0 ? (float)g->pitch_px : 0.0f;
        g->clump_roll_y = roll * time;
        g->clump_seed = g->clip_token ? g->clip_token : g->session_seed;
        if (g->clump_seed == 0) g->clump_seed = 1ULL;
        float sd = sd_spatial;
        g->clump_rms_norm = 1.0f;
        const float mix = clampf(g->clump_temporal_mix, 0.0f, 1.0f);
        if (mix > 0.0f) {
            const int period_i = g->clump_period <= 0 ? 1 : g->clump_period;
            const float period = (float)period_i;
            const float step_f = floorf(time / period);
            const int64_t step = (int64_t)step_f;
            const float frac = (time - step_f * period) / period;
            const float t = smoothstep_f(clampf(frac, 0.0f, 1.0f));
            const uint64_t ts = g->session_seed ? g->session_seed : 1ULL;
            const uint64_t sa = fjo_splitmix64(ts ^ ((uint64_t)step * 0xD2B74407B1CE6E93ULL));
            const uint64_t sb = fjo_splitmix64(ts ^ ((uint64_t)(step + 1) * 0xD2B74407B1CE6E93ULL));
            const float u1a = hash01(0, 0, sa ^ GOLDEN), u2a = hash01(1, 0, sa ^ MIX1);
            const float u1b = hash01(0, 0, sb ^ GOLDEN), u2b = hash01(1, 0, sb ^ MIX1);
            float u1 = u1a + (u1b - u1a) * t;
            float u2 = u2a + (u2b - u2a) * t;
            u1 = clampf(u1, 1e-6f, 1.0f - 1e-6f);
            u2 = clampf(u2, 0.0f, 1.0f);
            const float rt = sqrtf(-2.0f * fjo_det_logf(u1));
            const float nt = rt * fjo_det_cosf(two_pi * u2);
            const float strength_sd = mix * 0.23597824f;
            float strength = lognormal(1.0f, strength_sd, nt);
            strength *= 1.0f / sqrtf(1.0f + strength_sd * strength_sd);
            sd = sd_spatial * strength;
            const float den = 1.0f + sd * sd;
            const float num = 1.0f + sd_spatial * sd_spatial;
            g->clump_rms_norm = den > 0.0f ? sqrtf(num / den) : 1.0f;
        }
        const float sigma_sq = fjo_det_logf(1.0f + (sd * sd) / 1.0f);
        g->clump_sigma = sqrtf(fmaxf(0.0f, sigma_sq));
        g->clump_mu = fjo_det_logf(fmaxf(1e-12f, 1.0f)) - 0.5f * sigma_sq;
        g->clump_active = 1;
    }
}

static float breathing_factor(const Grain* g, uint64_t ax, uint64_t ay) {
    if (!g->breathing_active) return 1.0f;
    const float bx = (float)ax + g->breathing_drift_x;
    const float by = (float)ay + g->breathing_drift_y + g->breathing_roll_y;
    const float xs = bx / g->breathing_cell_small_px, ys = by / g->breathing_cell_small_px;
    const float xl = bx / g->breathing_cell_large_px, yl = by / g->breathing_cell_large_px;
    const float sa = value_noise(xs, ys, g->breathing_seed_a ^ GOLDEN);
    const float sb = value_noise(xs, ys, g->breathing_seed_b ^ GOLDEN);
    const float la = value_noise(xl, yl, g->breathing_seed_a ^ MIX1);
    const float lb = value_noise(xl, yl, g->breathing_seed_b ^ MIX1);
    const float ns = sa + (sb - sa) * g->breathing_t;
    const float nl = la + (lb - la) * g->breathing_t;
    const float noise = g->breathing_mix_u * ns + (1.0f - g->breathing_mix_u) * nl;
    const float n = noise * 2.0f - 1.0f;
    const float f = 1.0f + g->breathing_amp * n;
    return f > 0.0f ? f : 0.0f;
}

static float clump_factor(const Grain* g, uint64_t ax, uint64_t ay) {
    if (!g->clump_active) return 1.0f;
    const float xs = (float)ax / g->clump_cell_px;
    const float ys = ((float)ay + g->clump_roll_y) / g->clump_cell_px;
    const float u1s = value_noise(xs, ys, g->clump_seed ^ GOLDEN);
    const float u2s = value_noise(xs + 19.19f, ys + 7.23f, g->clump_seed ^ MIX1);
    const float u1 = clampf(u1s, 1e-6f, 1.0f - 1e-6f);
    const float u2 = clampf(u2s, 0.0f, 1.0f);
    const float r = sqrtf(-2.0f * fjo_det_logf(u1));
    const float n = r * fjo_det_cosf(6.28318530717958647692f * u2);
    float v = fjo_det_expf(g->clump_mu + g->clump_sigma * n);
    if (!is_finite_f(v)) v = 1.0f;
    return v * g->clump_rms_norm;
}

static float wang_sample(const Grain* g, int tile, int x, int y) {
    if (tile < 0 || tile >= g->wang_count) return 0.0f;
    const int w = g->wang_w, h = g->wang_h;
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x >= w) x = w - 1;
    if (y >= h) y = h - 1;
    const uint8_t v = g->wang_tiles[((size_t)tile * h + y) * w + x];
    return ((float)v + 0.5f) * (1.0f / 256.0f);
}

static void wang_offsets(const Grain* g, uint64_t ax, uint64_t ay, int* ox, int* oy) {
    *ox = *oy = 0;
    if (!(g->pixel_size_um > 0.0f) || !(g->wang_cell_mm > 0.0f)) return;
    const float xmm = (float)ax * (g->pixel_size_um * 0.001f);
    const float ymm = (float)ay * (g->pixel_size_um * 0.001f);
    const float inv = 1.0f / g->wang_cell_mm;
    const float cx = xmm * inv, cy = ymm * inv;
    const float fx = cx - floorf(cx), fy = cy - floorf(cy);
    const int64_t mx = (int64_t)floorf(cx), my = (int64_t)floorf(cy);
    const uint64_t s = g->clip_token ? g->clip_token : g->session_seed;
    const uint64_t base = s ? s : 1ULL;
    const uint32_t c = (uint32_t)g->wang_colors;
    const uint32_t l = hash_u32((int)mx, (int)my, base, 0xA5A5A5A5u) % c;
    const uint32_t r = hash_u32((int)(mx + 1), (int)my, base, 0x5A5A5A5Au) % c;
    const uint32_t t = hash_u32((int)mx, (int)my, base, 0xC3C3C3C3u) % c;
    const uint32_t b = hash_u32((int)mx, (int)(my + 1), base, 0x3C3C3C3Cu) % c;
    const int tile = g->wang_lut[(((size_t)l * c + r) * c + t) * c + b];
    const int w = g->wang_w, h = g->wang_h;
    int tx = (int)floorf(fx * (float)w), ty = (int)floorf(fy * (float)h);
    if (tx >= w) tx = w - 1;
    if (ty >= h) ty = h - 1;
    const float v0 = wang_sample(g, tile, tx, ty);
    const float v1 = wang_sample(g, tile, (tx + (w >> 1)) % w, (ty + (h >> 1)) % h);
    const int warp_x = g->stbn_w > 0 ? ((g->stbn_w / 8) > 4 ? g->stbn_w / 8 : 4) : 0;
    const int warp_y = g->stbn_h > 0 ? ((g->stbn_h / 8) > 4 ? g->stbn_h / 8 : 4) : 0;
    if (!warp_x || !warp_y) return;
    *ox = (int)floorf((v0 - 0.5f) * 2.0f * (float)warp_x);
    *oy = (int)floorf((v1 - 0.5f) * 2.0f * (float)warp_y);
}

/* One pixel's STBN stream (GrainRngDevice, …FilmPipeline.cu:1892-1995). */
typedef struct {
    const Grain* g;
    uint64_t seed, ax, ay;
    uint32_t draw;
    int wx, wy;
    int64_t pitch_offset;
    size_t row_base;
    int mask;
} Stream;

static void stream_init(Stream* s, const Grain* g, uint64_t seed, uint64_t ax, uint64_t ay, int frame_offset, int wx,
                        int wy) {
    s->g = g;
    s->seed = seed;
    s->ax = ax;
    s->ay = ay;
    s->draw = 0;
    s->wx = wx;
    s->wy = wy;
    s->mask = (g->stbn_w & (g->stbn_w - 1)) == 0 && (g->stbn_h & (g->stbn_h - 1)) == 0;
    s->pitch_offset = g->pitch_px > 0 ? (int64_t)g->pitch_px * (g->frame_index + (int64_t)frame_offset) : 0;
    const int frame = wrap_coord((int64_t)g->stbn_frame + frame_offset, g->stbn_frames);
    s->row_base = (size_t)frame * (size_t)g->stbn_h;
}

float fjo_stbn_uniform_at(const uint8_t* stbn, int w, int h, size_t row_base, int64_t xc, int64_t yc, int mask) {
    int x, y;
    if (mask) {
        x = (int)((uint64_t)xc & (uint64_t)(w - 1));
        y = (int)((uint64_t)yc & (uint64_t)(h - 1));
    } else {
        x = wrap_coord(xc, w);
        y = wrap_coord(yc, h);
    }
    return ((float)stbn[(row_base + (size_t)y) * (size_t)w + (size_t)x] + 0.5f) * (1.0f / 256.0f);
}

static float stream_uniform(Stream* s) {
    const Grain* g = s->g;
    const uint32_t draw = s->draw++;
    const uint64_t h = fjo_splitmix64(s->seed + (uint64_t)draw * GOLDEN);
    const int offx = (int)(h & 0xFFFFu) + s->wx;
    const int offy = (int)((h >> 16) & 0xFFFFu) + s->wy;
    const int64_t xc = (int64_t)s->ax + offx + g->stbn_off_x;
    int64_t yc = (int64_t)s->ay + offy + g->stbn_off_y;
    yc += s->pitch_offset;
    return fjo_stbn_uniform_at(g->stbn, g->stbn_w, g->stbn_h, s->row_base, xc, yc, s->mask);
}

static float stream_normal(Stream* s) {
    float u1 = stream_uniform(s);
    u1 = fminf(fmaxf(u1, 1e-7f), 1.0f);
    const float u2 = stream_uniform(s);
    const float r = sqrtf(-2.0f * fjo_det_logf(u1));
    return r * fjo_det_cosf(6.28318530717958647692f * u2);
}

static int poisson(float lambda, Stream* s) {
    if (!is_finite_f(lambda) || !(lambda > 0.0f)) return 0;
    if (lambda < 10.0f) {
        const float L = fjo_det_expf(-lambda);
        float p = 1.0f;
        int k = 0;
        while (p > L && k < 1024) {
            ++k;
            const float u = stream_uniform(s);
            p *= fminf(fmaxf(u, 1e-7f), 1.0f);
        }
        return k - 1;
    }
    const float z = stream_normal(s);
    float sample = lambda + sqrtf(lambda) * z;
    if (!is_finite_f(sample)) sample = 0.0f;
    int n = (int)floorf(sample + 0.5f);
    return n < 0 ? 0 : n;
}

static float particle_model(const Grain* g, float density, float dmax, float n_particles, float od, float uniformity,
                            uint64_t seed, uint64_t ax, uint64_t ay, int frame_offset, int wx, int wy) {
    if (!is_finite_f(density) || density < 0.0f) density = 0.0f;
    float p = density / dmax;
    p = fminf(fmaxf(p, 1e-6f), 1.0f - 1e-6f);
    const float u = is_finite_f(uniformity) ? fminf(fmaxf(uniformity, 0.0f), 1.0f) : 0.0f;
    float sat = 1.0f - p * u * (1.0f - 1e-6f);
    if (!is_finite_f(sat) || sat <= 0.0f) sat = 1e-6f;
    Stream s;
    stream_init(&s, g, seed, ax, ay, frame_offset, wx, wy);
    const float lambda = n_particles / sat;
    const int count = poisson(lambda * p, &s);
    float v = (float)count * od * sat;
    if (!is_finite_f(v)) v = 0.0f;
    return v;
}

static float interp_layer(float density, const float* x, const float* y, int n, int positive, int axis_finite,
                          const float* prefix) {
    if (!x || !y || n <= 0) return density;
    if (!is_finite_f(density)) return y[0];
    const float q = positive ? -density : density;
#define AX(i) (positive ? -x[(i)] : x[(i)])
    const int bounded = axis_finite && prefix && n == 256;
    int begin = 0, end = n - 1;
    if (!bounded) {
        while (begin < n && !is_finite_f(AX(begin))) ++begin;
        if (begin >= n) return density;
        while (end > begin && !is_finite_f(AX(end))) --end;
    }
    const float xmin = AX(begin), xmax = AX(end);
    if (!is_finite_f(xmin) || !is_finite_f(xmax) || !(xmax >= xmin)) return y[begin];
    if (q <= xmin) return y[begin];
    if (q >= xmax) return y[end];
    int i1 = begin + 1;
    if (bounded) {
        int bb = 0, be = 15;
        for (int step = 0; step < 4; ++step) {
            const int mid = bb + (be - bb) / 2;
            if (prefix[mid] < q)
                bb = mid + 1;
            else
                be = mid;
        }
        i1 = bb * 16 > 1 ? bb * 16 : 1;
        const int search_end = end < (bb + 1) * 16 - 1 ? end : (bb + 1) * 16 - 1;
        while (i1 <= search_end && AX(i1) < q) ++i1;
    } else {
        while (i1 <= end && AX(i1) < q) ++i1;
    }
    if (i1 > end) return y[end];
    const int i0 = i1 - 1;
    const float x0 = AX(i0), x1 = AX(i1), y0 = y[i0], y1 = y[i1];
#undef AX
    const float den = x1 - x0;
    if (!(den > 0.0f) || !is_finite_f(den)) return y0;
    const float t = (q - x0) / den;
    return y0 + t * (y1 - y0);
}

/* Particle samples of one channel at one pixel for a pass (fine-only inside a
 * pass, so the size mix never splits a sample). */
static float retime(const Grain* g, float density, float s0, float s1, float bf, float cf, int want_next) {
    float a = density + (s0 - density) * bf * cf;
    if (want_next) {
        const float b = density + (s1 - density) * bf * cf;
        const float d0 = a - density, d1 = b - density;
        const float blend = d0 + (d1 - d0) * g->time_alpha;
        const float w0 = 1.0f - g->time_alpha;
        const float den = w0 * w0 + g->time_alpha * g->time_alpha + 1e-6f;
        a = density + blend * (1.0f / sqrtf(den));
    }
    return a;
}

static void pixel_setup(const Grain* g, int x, int y, uint64_t* ax, uint64_t* ay, int* wx, int* wy) {
    *ax = (uint64_t)((int64_t)g->origin_x + x);
    *ay = (uint64_t)((int64_t)g->origin_y + y);
    wang_offsets(g, *ax, *ay, wx, wy);
}

static float layer_sample(const Grain* g, float in_density, int c, int layer, uint64_t ax, uint64_t ay, int wx, int wy,
                          float bf, float cf) {
    const int want_next = g->time_alpha > 1e-6f && g->time_alpha < 0.999999f;
    float density = interp_layer(in_density, g->base_curve[c], g->layer_curve[layer][c], g->curve_n, g->positive_film,
                                 g->axis_finite[c], g->axis_prefix[c]);
    density += g->dmin_l[layer][c];
    const uint64_t salt = (uint64_t)c + (uint64_t)layer * 10ULL;
    const float dmax = g->dmax_l[layer][c], n = g->n_l[layer][c], od = g->od_l[layer][c], u = g->uniformity[c];
    const float s0 = particle_model(g, density, dmax, n, od, u, g->seed_base ^ salt, ax, ay, 0, wx, wy);
    const float s1 = want_next ? particle_model(g, density, dmax, n, od, u, g->seed_base_next ^ salt, ax, ay, 1, wx, wy)
                               : 0.0f;
    return retime(g, density, s0, s1, bf, cf, want_next);
}

static float simple_sample(const Grain* g, float in_density, int c, uint64_t ax, uint64_t ay, int wx, int wy,
                           float bf, float cf) {
    const int want_next = g->time_alpha > 1e-6f && g->time_alpha < 0.999999f;
    const float density = in_density + g->density_min[c];
    const int nsub = g->n_sublayers > 0 ? g->n_sublayers : 1;
    float acc = 0.0f, acc_next = 0.0f;
    for (int sl = 0; sl < nsub; ++sl) {
        const uint64_t salt = (uint64_t)c + (uint64_t)sl * 10ULL;
        acc += particle_model(g, density, g->density_max[c], g->n_particles[c], g->od_particle[c], g->uniformity[c],
                              g->seed_base ^ salt, ax, ay, 0, wx, wy);
        if (want_next)
            acc_next += particle_model(g, density, g->density_max[c], g->n_particles[c], g->od_particle[c],
                                       g->uniformity[c], g->seed_base_next ^ salt, ax, ay, 1, wx, wy);
    }
    acc /= (float)nsub;
    if (want_next) acc_next /= (float)nsub;
    return retime(g, density, acc, acc_next, bf, cf, want_next);
}

/* ---- separable blur (half-sample reflect) ---------------------------------- */
static int reflect_half_wide(int64_t i, int n) {
    if (n <= 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * (int64_t)n - i - 1;
    return (int)i;
}

static void blur_h(const float* in, float* out, int W, int H, const Fir* k) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (int j = 0; j <= 2 * k->r; ++j)
                acc = fmaf(in[(size_t)y * W + reflect_half_wide((int64_t)x + j - k->r, W)], k->w[j], acc);
            out[(size_t)y * W + x] = is_finite_f(acc) ? acc : 0.0f;
        }
}

/* Vertical pass accumulated into dst: dst = prev + [weight *] blur; with
 * `sub`, then (dst + bias) - sub (grain_blur_vertical_accumulate, …ScanPipeline.cu:457-545). */
static void blur_v_acc(const float* in, float* dst, int W, int H, const Fir* k, int weighted, float weight, int init,
                       const float* sub, float bias) {
    const float wf = is_finite_f(weight) ? weight : 0.0f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (int j = 0; j <= 2 * k->r; ++j)
                acc = fmaf(in[(size_t)reflect_half_wide((int64_t)y + j - k->r, H) * W + x], k->w[j], acc);
            const float blurred = is_finite_f(acc) ? acc : 0.0f;
            const size_t i = (size_t)y * W + x;
            const float prev = init ? 0.0f : dst[i];
            float v = weighted ? prev + wf * blurred : prev + blurred;
            v = is_finite_f(v) ? v : 0.0f;
            if (!weighted && sub) {
                if (bias != 0.0f) {
                    const float b = v + bias;
                    v = is_finite_f(b) ? b : 0.0f;
                }
                const float d = v - sub[i];
                v = is_finite_f(d) ? d : 0.0f;
            }
            dst[i] = v;
        }
}

static float san(float v) { return is_finite_f(v) ? v : 0.0f; }

/* ---- schedule -------------------------------------------------------------- */
typedef struct {
    const FjPreparedHostData* p;
    Grain base;
    Fir corr[3];
    Fir dye[3][3];
    int W, H;
    size_t N;
    float *tmp, *scale, *delta, *layer, *shared;
} Ctx;

static void layer_planes(Ctx* cx, const Grain* g, const float* dens, int c, int mode, float* out) {
    /* mode: -1 triplet sum, 0..2 one layer, 3 simple (no sub-layers) */
    for (int y = 0; y < cx->H; ++y)
        for (int x = 0; x < cx->W; ++x) {
            const size_t i = (size_t)y * cx->W + x;
            uint64_t ax, ay;
            int wx, wy;
            pixel_setup(g, x, y, &ax, &ay, &wx, &wy);
            const float cf = clump_factor(g, ax, ay);
            const float bf = breathing_factor(g, ax, ay);
            if (mode == 3) {
                out[i] = san(simple_sample(g, dens[i], c, ax, ay, wx, wy, bf, cf));
            } else if (mode < 0) {
                float sum = 0.0f;
                for (int l = 0; l < 3; ++l) sum += san(layer_sample(g, dens[i], c, l, ax, ay, wx, wy, bf, cf));
                out[i] = san(sum);
            } else {
                out[i] = san(layer_sample(g, dens[i], c, mode, ax, ay, wx, wy, bf, cf));
            }
        }
}

static int generate_scale(Ctx* cx, const Grain* base, const float* dens, int c, uint64_t salt, float factor,
                          const Fir* corr, int apply_corr, float weight, int init) {
    if (!is_finite_f(factor) || !(factor > 0.0f)) return 2;
    Grain g = *base;
    g.seed_base ^= salt;
    g.seed_base_next ^= salt;
    g.size_mix_weight = 0.0f;
    g.size_mix_scale = 1.0f;
    for (int k = 0; k < 3; ++k) {
        g.n_particles[k] = base->n_particles[k] / factor;
        g.od_particle[k] = base->od_particle[k] * factor;
        for (int l = 0; l < 3; ++l) {
            g.n_l[l][k] = base->n_l[l][k] / factor;
            g.od_l[l][k] = base->od_l[l][k] * factor;
        }
    }
    const size_t N = cx->N;
    const float bias = -base->density_min[c];
    int formed = 0;
    if (g.sublayers) {
        const Fir* d0 = &cx->dye[0][c];
        const Fir* d1 = &cx->dye[1][c];
        const Fir* d2 = &cx->dye[2][c];
        /* Equal (radius, sigma) share one prepared kernel: sum the layers, blur once. */
        const int same = d0->r == d1->r && d0->r == d2->r &&
                         (d0->r == 0 || (d0->sigma == d1->sigma && d0->sigma == d2->sigma));
        if (same) {
            layer_planes(cx, &g, dens, c, -1, cx->layer);
            if (d0->r > 0) {
                blur_h(cx->layer, cx->tmp, cx->W, cx->H, d0);
                blur_v_acc(cx->tmp, cx->scale, cx->W, cx->H, d0, 0, 0.0f, 1, dens, bias);
                formed = 1;
            } else {
                for (size_t i = 0; i < N; ++i) cx->scale[i] = san(0.0f + cx->layer[i]);
            }
        } else {
            for (int l = 0; l < 3; ++l) {
                layer_planes(cx, &g, dens, c, l, cx->layer);
                const Fir* d = &cx->dye[l][c];
                if (d->r > 0) {
                    blur_h(cx->layer, cx->tmp, cx->W, cx->H, d);
                    blur_v_acc(cx->tmp, cx->scale, cx->W, cx->H, d, 0, 0.0f, l == 0, l == 2 ? dens : NULL,
                               l == 2 ? bias : 0.0f);
                    formed = l == 2;
                } else {
                    for (size_t i = 0; i < N; ++i) cx->scale[i] = san((l == 0 ? 0.0f : cx->scale[i]) + cx->layer[i]);
                }
            }
        }
    } else {
        layer_planes(cx, &g, dens, c, 3, cx->scale);
    }
    if (!formed) {
        if (bias != 0.0f) {
            for (size_t i = 0; i < N; ++i) {
                const float b = san(cx->scale[i] + bias);
                cx->scale[i] = san((b - dens[i]) * 1.0f);
            }
        } else {
            for (size_t i = 0; i < N; ++i) cx->scale[i] = san((cx->scale[i] - dens[i]) * 1.0f);
        }
    }
    if (apply_corr && corr->r > 0) {
        blur_h(cx->scale, cx->tmp, cx->W, cx->H, corr);
        blur_v_acc(cx->tmp, cx->delta, cx->W, cx->H, corr, 1, weight, init, NULL, 0.0f);
        return 0;
    }
    const float w = is_finite_f(weight) ? weight : 0.0f;
    for (size_t i = 0; i < N; ++i) cx->delta[i] = san((init ? 0.0f : cx->delta[i]) + w * cx->scale[i]);
    return 0;
}

static int generate_mix(Ctx* cx, const Grain* base, const float* dens, int c) {
    const FjVisualGrain* vg = &cx->p->grain;
    const float w_fine = vg->size_mix_weights[0], w_mid = vg->size_mix_weights[1], w_coarse = vg->size_mix_weights[2];
    int init = 1, rc;
    if (w_fine > 0.0f) {
        if ((rc = generate_scale(cx, base, dens, c, 0xA24BAED4963EE407ULL, 1.0f, &cx->corr[0], cx->corr[0].r > 0, w_fine,
                                 init)))
            return rc;
        init = 0;
    }
    if (w_mid > 0.0f) {
        if ((rc = generate_scale(cx, base, dens, c, 0x9F6C1E6B2C4D7A13ULL, sqrtf(vg->size_mix_scale), &cx->corr[1], 1,
                                 w_mid, init)))
            return rc;
        init = 0;
    }
    if (w_coarse > 0.0f) {
        if ((rc = generate_scale(cx, base, dens, c, 0x3C79AC492BA7B653ULL, vg->size_mix_scale, &cx->corr[2], 1, w_coarse,
                                 init)))
            return rc;
        init = 0;
    }
    const float gain = is_finite_f(vg->size_mix_gain) ? vg->size_mix_gain : 0.0f;
    for (size_t i = 0; i < cx->N; ++i) cx->delta[i] = san(cx->delta[i] * gain);
    return 0;
}

static void debug_avg3(float* const d[3], size_t n, float offset, float scale) {
    const float s = is_finite_f(scale) ? scale : 1.0f, o = is_finite_f(offset) ? offset : 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float a = (d[0][i] + d[1][i] + d[2][i]) * (1.0f / 3.0f);
        if (!is_finite_f(a)) a = 0.0f;
        d[0][i] = d[1][i] = d[2][i] = o + a * s;
    }
}

static int load_fir(const FjGrainGaussian* gg, Fir* out) {
    out->w = NULL;
    out->r = 0;
    out->sigma = 0.0f;
    if (!(isfinite(gg->sigma_px) && gg->sigma_px > 0.0f) || gg->radius <= 0) return 0;
    out->sigma = gg->sigma_px;
    if (gg->radius > 4096) return 2;
    out->w = (float*)malloc(sizeof(float) * (size_t)(2 * gg->radius + 1));
    if (!out->w || fjo_gaussian_weights(gg->radius, gg->sigma_px, out->w)) return 2;
    out->r = gg->radius;
    return 0;
}

static int build_grain(const FjPreparedHostData* p, Grain* g) {
    const FjVisualGrain* vg = &p->grain;
    const FjStaticNoise* nz = &p->noise;
    memset(g, 0, sizeof(*g));
    g->stbn = nz->stbn.data;
    g->stbn_w = nz->stbn_width;
    g->stbn_h = nz->stbn_height;
    g->stbn_frames = nz->stbn_frames;
    if (!g->stbn || g->stbn_w <= 0 || g->stbn_h <= 0 || g->stbn_frames <= 0) return 2;
    if (!nz->wang_tiles.data || !nz->wang_lut.data || nz->wang_width <= 0 || nz->wang_height <= 0 ||
        nz->wang_tile_count == 0 || nz->wang_colors <= 0)
        return 2;
    const uint64_t ox[2] = {vg->session_seed, 0xA5u}, oy[2] = {vg->session_seed, 0x5Au};
    g->stbn_off_x = (int)(fnv_u64s(ox, 2) % (uint64_t)g->stbn_w);
    g->stbn_off_y = (int)(fnv_u64s(oy, 2) % (uint64_t)g->stbn_h);
    {
        const int64_t phase = (int64_t)(vg->session_seed % (uint64_t)g->stbn_frames);
        int f = (int)((vg->frame0 + phase) % g->stbn_frames);
        if (f < 0) f += g->stbn_frames;
        g->stbn_frame = f;
    }
    g->frame_index = vg->frame0;
    g->wang_tiles = nz->wang_tiles.data;
    g->wang_lut = nz->wang_lut.data;
    g->wang_w = nz->wang_width;
    g->wang_h = nz->wang_height;
    g->wang_count = (int)nz->wang_tile_count;
    g->wang_colors = nz->wang_colors;
    g->wang_cell_mm = vg->wang_cell_mm;
    g->pixel_size_um = vg->pixel_size_um;
    g->session_seed = vg->session_seed;
    g->clip_token = vg->clip_token;
    g->origin_x = vg->render_extent.x;
    g->origin_y = vg->render_extent.y;
    g->pitch_px = vg->pitch_px;
    g->time_alpha = vg->frame_alpha;
    g->seed_base = vg->seed_base;
    g->seed_base_next = vg->seed_base_next;
    g->sublayers = (vg->flags & FJ_GRAIN_SUBLAYERS) != 0;
    g->positive_film = p->capture_polarity == FJ_POLARITY_POSITIVE;
    g->n_sublayers = (int)vg->sublayer_count;
    const FjFilmDevelopment* fd = &p->film_development;
    g->curve_n = (int)fd->log_exposure.count;
    for (int c = 0; c < 3; ++c) {
        g->density_min[c] = vg->particle_density_min_cmy[c];
        g->uniformity[c] = vg->uniformity_cmy[c];
        g->density_max[c] = vg->density_max_cmy[c];
        g->n_particles[c] = vg->n_particles_cmy[c];
        g->od_particle[c] = vg->od_particle_cmy[c];
        g->axis_finite[c] = g->sublayers && (vg->flags & (1u << (3u + c))) != 0;
        for (int b = 0; b < 16; ++b) g->axis_prefix[c][b] = vg->layer_axis_block_prefix_max[c][b];
        for (int l = 0; l < 3; ++l) {
            g->dmin_l[l][c] = vg->density_min_layers[l][c];
            g->dmax_l[l][c] = vg->density_max_layers[l][c];
            g->n_l[l][c] = vg->n_particles_layers[l][c];
            g->od_l[l][c] = vg->od_particle_layers[l][c];
            g->layer_curve[l][c] = g->sublayers ? fd->density_layers[l][c].data : NULL;
        }
    }
    g->breathing_period = vg->breathing_period_frames;
    g->clump_period = vg->clump_morph_period_frames;
    g->breathing_amp = vg->breathing_amplitude;
    g->breathing_cell_small_um = vg->breathing_cell_um_small;
    g->breathing_cell_large_um = vg->breathing_cell_um_large;
    g->breathing_mix = vg->breathing_mix;
    g->breathing_drift = vg->breathing_drift_um_per_frame;
    g->micro[0] = vg->micro_structure[0];
    g->micro[1] = vg->micro_structure[1];
    g->clump_temporal_mix = vg->clump_temporal_mix;
    frame_uniforms(g);
    return 0;
}

int fjo_grain_apply(const FjPreparedHostData* p, float* const dens[3], int W, int H) {
    if (!p || !(p->grain.flags & FJ_GRAIN_ACTIVE) || W <= 0 || H <= 0) return 2;
    const FjVisualGrain* vg = &p->grain;
    Ctx cx;
    memset(&cx, 0, sizeof(cx));
    cx.p = p;
    cx.W = W;
    cx.H = H;
    cx.N = (size_t)W * H;
    int rc = build_grain(p, &cx.base);
    float* base_curves = NULL;
    if (!rc && cx.base.sublayers) {
        const int n = cx.base.curve_n;
        base_curves = (float*)malloc(sizeof(float) * 3 * (size_t)n);
        if (!base_curves) rc = 2;
        for (int c = 0; !rc && c < 3; ++c) {
            for (int i = 0; i < n; ++i) base_curves[c * n + i] = p->film_development.density_rgb.data[i * 3 + c];
            cx.base.base_curve[c] = base_curves + (size_t)c * n;
        }
    }
    for (int k = 0; !rc && k < 3; ++k) rc = load_fir(&vg->correlation[k], &cx.corr[k]);
    for (int l = 0; !rc && l < 3; ++l)
        for (int c = 0; !rc && c < 3; ++c) rc = load_fir(&vg->dye_cloud[l][c], &cx.dye[l][c]);
    float* buf = rc ? NULL : (float*)calloc(cx.N * 5, sizeof(float));
    if (!rc && !buf) rc = 2;
    if (!rc) {
        cx.tmp = buf;
        cx.scale = buf + cx.N;
        cx.delta = buf + 2 * cx.N;
        cx.layer = buf + 3 * cx.N;
        cx.shared = buf + 4 * cx.N;
        const int debug = (int)vg->debug_view;
        const float amplitude = vg->amplitude;
        if (debug == 6) {
            debug_avg3(dens, cx.N, 0.0f, vg->debug_scale * 4.0f);
        } else if (debug >= 2 && debug <= 5) {
            const int coarse = debug == 3 || debug == 5, raw = debug == 4 || debug == 5;
            for (int c = 0; !rc && c < 3; ++c) {
                if (!coarse || vg->size_mix_weights[2] > 0.0f) {
                    const Fir* k = coarse ? &cx.corr[2] : &cx.corr[0];
                    rc = generate_scale(&cx, &cx.base, dens[c], c, coarse ? 0x3C79AC492BA7B653ULL : 0xA24BAED4963EE407ULL,
                                        coarse ? vg->size_mix_scale : 1.0f, k, !raw && k->r > 0, 1.0f, 1);
                } else {
                    memset(cx.delta, 0, sizeof(float) * cx.N);
                }
                const float s = is_finite_f(amplitude) ? amplitude : 0.0f;
                for (size_t i = 0; i < cx.N; ++i) dens[c][i] = san(0.0f + s * cx.delta[i]);
            }
            if (!rc) debug_avg3(dens, cx.N, 0.5f, vg->debug_scale);
        } else {
            const int shared = (debug == 0 || debug == 1) && vg->chroma_mix < 0.999f && vg->chroma_shared_weight > 0.0f;
            if (shared) {
                Grain sg = cx.base;
                sg.seed_base ^= 0x7E8A1B9D3F2C65A1ULL;
                sg.seed_base_next ^= 0x7E8A1B9D3F2C65A1ULL;
                rc = generate_mix(&cx, &sg, dens[1], 1);
                if (!rc) memcpy(cx.shared, cx.delta, sizeof(float) * cx.N);
            }
            for (int c = 0; !rc && c < 3; ++c) {
                if (vg->chroma_independent_weight > 0.0f || !shared)
                    rc = generate_mix(&cx, &cx.base, dens[c], c);
                else
                    memset(cx.delta, 0, sizeof(float) * cx.N);
                if (rc) break;
                if (shared) {
                    const float sw = clampf(vg->chroma_shared_weight, 0.0f, 1.0f);
                    const float iw = clampf(vg->chroma_independent_weight, 0.0f, 1.0f);
                    const int ind = vg->chroma_independent_weight > 0.0f;
                    for (size_t i = 0; i < cx.N; ++i) {
                        const float s = sw > 0.0f ? cx.shared[i] : 0.0f;
                        const float d = (ind && iw > 0.0f) ? cx.delta[i] : 0.0f;
                        cx.delta[i] = san(amplitude * (sw * s + iw * d));
                    }
                }
                const float sc = shared ? 1.0f : amplitude;
                const float s = is_finite_f(sc) ? sc : 0.0f;
                for (size_t i = 0; i < cx.N; ++i) dens[c][i] = san((debug == 0 ? dens[c][i] : 0.0f) + s * cx.delta[i]);
            }
            if (!rc && debug != 0) debug_avg3(dens, cx.N, 0.5f, vg->debug_scale);
        }
    }
    free(buf);
    free(base_curves);
    for (int k = 0; k < 3; ++k) free(cx.corr[k].w);
    for (int l = 0; l < 3; ++l)
        for (int c = 0; c < 3; ++c) free(cx.dye[l][c].w);
    return rc;
}

/* One uniform of a pixel's stream, for stream-level tests. */
float fjo_grain_uniform(const FjPreparedHostData* p, uint64_t seed, int x, int y, int frame_offset, uint32_t draw) {
    Grain g;
    if (build_grain(p, &g)) return NAN;
    uint64_t ax, ay;
    int wx, wy;
    pixel_setup(&g, x, y, &ax, &ay, &wx, &wy);
    Stream s;
    stream_init(&s, &g, seed, ax, ay, frame_offset, wx, wy);
    s.draw = draw;
    return stream_uniform(&s);
}
