// fj_grain_host.cpp — visual grain: descriptor rules, payload, frame uniforms
// and the stage schedule.
//
// The schedule restates juicer_cuda_apply_visual_grain
// (src/Cuda/Scan/JuicerCudaScanPipeline.cu:3141-3756): per channel, the fine,
// mid and coarse scales (seed salts, particle-count factors 1, sqrt(scale),
// scale), each scale's sub-layer samples (summed before one shared dye-cloud
// blur when all three layers carry the same kernel), its delta against the
// developed density, the correlation blur weighted into the channel delta,
// the size-mix gain, the shared-chroma mix and the reconstruction. Sample
// planes whose delta needs no dye-cloud blur are formed inside the sample
// kernel.
#include "fj_grain_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "fj_detmath.h"
#include "fj_host.h"
#include "fj_spatial.h"

#pragma clang fp contract(off)

namespace fjx {

namespace {

constexpr uint64_t kSaltFine = 0xA24BAED4963EE407ull;
constexpr uint64_t kSaltMid = 0x9F6C1E6B2C4D7A13ull;
constexpr uint64_t kSaltCoarse = 0x3C79AC492BA7B653ull;
constexpr uint64_t kSaltShared = 0x7E8A1B9D3F2C65A1ull;
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kMix1 = 0xBF58476D1CE4E5B9ull;
constexpr int kMaxGrainRadius = 75;  // the frame plan's Gaussian cap (film/grain/frame.rs:414-418)

uint64_t fnv_u64s(std::initializer_list<uint64_t> vals) {
    uint64_t h = 0xCBF29CE484222325ull;
    for (uint64_t v : vals)
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xFFu;
            h *= 0x100000001B3ull;
        }
    return h;
}

float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
float clampf(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }

float hash01(int x, int y, uint64_t seed) {
    const uint64_t h = splitmix64(seed ^ (static_cast<uint64_t>(static_cast<int64_t>(x)) * kGolden) ^
                                  (static_cast<uint64_t>(static_cast<int64_t>(y)) * kMix1));
    return static_cast<float>(static_cast<uint32_t>(h)) * (1.0f / 4294967296.0f);
}

bool finite_positive(float v) { return std::isfinite(v) && v > 0.0f; }

// Sparse content identity of a large static byte span (its address, length
// and 4096 sampled bytes): the static noise is uploaded once per owner.
uint64_t span_identity(const FjByteSpan& s, uint64_t tag) {
    uint64_t h = hash_bytes(&s.data, sizeof(s.data), tag);
    h = hash_bytes(&s.count, sizeof(s.count), h);
    if (s.count == 0 || !s.data) return h;
    const size_t step = std::max<size_t>(1, s.count / 4096);
    std::vector<uint8_t> sample;
    for (size_t i = 0; i < s.count; i += step) sample.push_back(s.data[i]);
    sample.push_back(s.data[s.count - 1]);
    return hash_bytes(sample.data(), sample.size(), h);
}

bool gaussian_ready(const FjGrainGaussian& g) {
    if (g.radius <= 0) return g.hash == 0;
    return g.hash != 0 && g.radius <= kMaxGrainRadius && std::isfinite(g.sigma_px) && g.sigma_px > 0.0f;
}

}  // namespace

bool grain_decode(const FjVisualGrain& g, std::string& field) {
    constexpr uint32_t kFlags = FJ_GRAIN_ACTIVE | FJ_GRAIN_SUBLAYERS | FJ_GRAIN_FULL_FRAME | FJ_GRAIN_AXIS_FINITE_C |
                                FJ_GRAIN_AXIS_FINITE_M | FJ_GRAIN_AXIS_FINITE_Y;
    if ((g.flags & ~kFlags) != 0 || g.debug_view > FJ_GRAIN_DEBUG_MEAN_DENSITY ||
        g.sublayer_count > static_cast<size_t>(std::numeric_limits<int>::max()))
        return field = "grain.flags_or_count", false;
    if ((g.flags & FJ_GRAIN_ACTIVE) == 0) {
        if (g.flags == 0 && g.hash == 0 && g.recipe_hash == 0) return true;
        return field = "grain.inactive", false;
    }
    if (g.hash == 0 || g.recipe_hash == 0) return field = "grain.identity", false;
    return true;
}

bool grain_binding(const FjVisualGrain& g, const FjFrame& f) {
    if (g.hash == 0) return true;
    auto same = [](const FjExtent& e, const FjRect& r) {
        return e.x == r.x1 && e.y == r.y1 && e.width == r.x2 - r.x1 && e.height == r.y2 - r.y1;
    };
    const double fl = std::floor(f.time_frames);
    return same(g.render_extent, f.render_window) && same(g.full_frame_extent, f.full_frame_extent) &&
           g.session_seed == f.session_seed && g.clip_token == f.clip_token && g.pixel_size_um == f.pixel_size_um &&
           g.frame0 == static_cast<int64_t>(fl) &&
           g.frame_alpha == static_cast<float>(std::clamp(f.time_frames - fl, 0.0, 1.0));
}

void grain_frame_uniforms(const FjVisualGrain& vg, GrainDev& g) {
    const float time = static_cast<float>(vg.frame0) + vg.frame_alpha;
    const float two_pi = 6.28318530717958647692f;
    g.breathing_active = 0;
    g.breathing_amp = vg.breathing_amplitude;
    if (vg.breathing_amplitude > 0.0f && vg.breathing_period_frames > 0 && vg.pixel_size_um > 0.0f &&
        vg.breathing_cell_um_small > 0.0f && vg.breathing_cell_um_large > 0.0f) {
        const float period = static_cast<float>(vg.breathing_period_frames);
        const float step_f = std::floor(time / period);
        const int64_t step = static_cast<int64_t>(step_f);
        const float frac = (time - step_f * period) / period;
        g.breathing_t = smooth(clampf(frac, 0.0f, 1.0f));
        const uint64_t seed = vg.session_seed ? vg.session_seed : 1ull;
        g.breathing_seed_a = splitmix64(seed ^ (static_cast<uint64_t>(step) * 0x8EBC6AF09C88C6E3ull));
        g.breathing_seed_b = splitmix64(seed ^ (static_cast<uint64_t>(step + 1) * 0x8EBC6AF09C88C6E3ull));
        const float inv_pixel = 1.0f / vg.pixel_size_um;
        g.breathing_cell_small_px = std::fmax(4.0f, vg.breathing_cell_um_small * inv_pixel);
        g.breathing_cell_large_px = std::fmax(4.0f, vg.breathing_cell_um_large * inv_pixel);
        float drift_px = 0.0f;
        if (vg.breathing_drift_um_per_frame > 0.0f) drift_px = vg.breathing_drift_um_per_frame * inv_pixel;
        float dx = 0.0f, dy = 0.0f;
        if (drift_px > 0.0f) {
            const uint64_t h = splitmix64(seed ^ 0x6A09E667F3BCC909ull);
            const float angle = static_cast<float>(static_cast<uint32_t>(h)) * (1.0f / 4294967296.0f) * two_pi;
            dx = det_cosf(angle) * drift_px;
            dy = det_sinf(angle) * drift_px;
        }
        const float roll = vg.pitch_px > 0 ? static_cast<float>(vg.pitch_px) : 0.0f;
        g.breathing_drift_x = dx * time;
        g.breathing_drift_y = dy * time;
        g.breathing_roll_y = roll * time;
        g.breathing_mix = clampf(vg.breathing_mix, 0.0f, 1.0f);
        g.breathing_active = 1;
    }
    g.clump_active = 0;
    const float sd_spatial = vg.micro_structure[1] * 0.001f;
    if (sd_spatial > 0.0f && vg.pixel_size_um > 0.0f) {
        float cell_um = vg.micro_structure[0];
        if (!(cell_um > 0.0f)) cell_um = 0.0f;
        const float inv_pixel = 1.0f / vg.pixel_size_um;
        g.clump_cell_px = cell_um * inv_pixel;
        if (!(g.clump_cell_px > 4.0f)) g.clump_cell_px = 4.0f;
        const float roll = vg.pitch_px > 0 ? static_cast<float>(vg.pitch_px) : 0.0f;
        g.clump_roll_y = roll * time;
        g.clump_seed = vg.clip_token ? vg.clip_token : vg.session_seed;
        if (g.clump_seed == 0) g.clump_seed = 1ull;
        float sd = sd_spatial;
        g.clump_rms_norm = 1.0f;
        const float mix = clampf(vg.clump_temporal_mix, 0.0f, 1.0f);
        if (mix > 0.0f) {
            const int period_i = vg.clump_morph_period_frames <= 0 ? 1 : vg.clump_morph_period_frames;
            const float period = static_cast<float>(period_i);
            const float step_f = std::floor(time / period);
            const int64_t step = static_cast<int64_t>(step_f);
            const float frac = (time - step_f * period) / period;
            const float t = smooth(clampf(frac, 0.0f, 1.0f));
            const uint64_t ts = vg.session_seed ? vg.session_seed : 1ull;
            const uint64_t sa = splitmix64(ts ^ (static_cast<uint64_t>(step) * 0xD2B74407B1CE6E93ull));
            const uint64_t sb = splitmix64(ts ^ (static_cast<uint64_t>(step + 1) * 0xD2B74407B1CE6E93ull));
            const float u1a = hash01(0, 0, sa ^ kGolden), u2a = hash01(1, 0, sa ^ kMix1);
            const float u1b = hash01(0, 0, sb ^ kGolden), u2b = hash01(1, 0, sb ^ kMix1);
            float u1 = u1a + (u1b - u1a) * t;
            float u2 = u2a + (u2b - u2a) * t;
            u1 = clampf(u1, 1e-6f, 1.0f - 1e-6f);
            u2 = clampf(u2, 0.0f, 1.0f);
            const float rt = std::sqrt(-2.0f * det_logf(u1));
            const float nt = rt * det_cosf(two_pi * u2);
            const float strength_sd = mix * 0.23597824f;
            // lognormal_from_mean_std_device with mean 1
            const float sigma_sq_s = det_logf(1.0f + (strength_sd * strength_sd / 1.0f));
            const float sigma_s = std::sqrt(std::fmax(0.0f, sigma_sq_s));
            const float mu_s = det_logf(std::fmax(1e-12f, 1.0f)) - 0.5f * sigma_sq_s;
            float strength = det_expf(mu_s + sigma_s * nt);
            strength *= 1.0f / std::sqrt(1.0f + strength_sd * strength_sd);
            sd = sd_spatial * strength;
            const float den = 1.0f + sd * sd;
            const float num = 1.0f + sd_spatial * sd_spatial;
            g.clump_rms_norm = den > 0.0f ? std::sqrt(num / den) : 1.0f;
        }
        const float sigma_sq = det_logf(1.0f + (sd * sd) / 1.0f);
        g.clump_sigma = std::sqrt(std::fmax(0.0f, sigma_sq));
        g.clump_mu = det_logf(std::fmax(1e-12f, 1.0f)) - 0.5f * sigma_sq;
        g.clump_active = 1;
    }
}

hipError_t grain_prepare(DeviceState& dev, const FjPreparedHostData& p, GrainPlan& out, bool& ok, std::string& diag,
                         bool& launch_rejected) {
    ok = false;
    launch_rejected = false;
    out = GrainPlan{};
    const FjVisualGrain& vg = p.grain;
    const FjStaticNoise& nz = p.noise;
    const bool sub = (vg.flags & FJ_GRAIN_SUBLAYERS) != 0;
    // pack_visual_grain_payload
    if (vg.render_extent.width <= 0 || vg.render_extent.height <= 0 || !(vg.pixel_size_um > 0.0f) ||
        vg.pitch_px <= 0 || p.capture_polarity > FJ_POLARITY_POSITIVE) {
        diag = "ResourceDescriptorMismatch phase=grain_payload field=descriptor";
        return hipSuccess;
    }
    if (!nz.stbn.data || nz.stbn_width <= 0 || nz.stbn_height <= 0 || nz.stbn_frames <= 0) {
        diag = "MissingRequiredResource phase=grain_payload field=stbn";
        return hipSuccess;
    }
    if (!nz.wang_tiles.data || !nz.wang_lut.data || nz.wang_width <= 0 || nz.wang_height <= 0 ||
        nz.wang_tile_count == 0 || nz.wang_colors <= 0) {
        diag = "ResourceDescriptorMismatch phase=grain_payload field=wang_identity";
        return hipSuccess;
    }
    for (const FjGrainGaussian& g : vg.correlation)
        if (!gaussian_ready(g)) {
            diag = "ResourceDescriptorMismatch phase=grain_payload field=correlation_gaussian";
            return hipSuccess;
        }
    for (const auto& row : vg.dye_cloud)
        for (const FjGrainGaussian& g : row)
            if (!gaussian_ready(g)) {
                diag = "ResourceDescriptorMismatch phase=grain_payload field=dye_cloud_gaussian";
                return hipSuccess;
            }
    const FjFilmDevelopment& fd = p.film_development;
    const int n = static_cast<int>(fd.log_exposure.count);
    if (sub) {
        if (vg.density_layers_hash == 0 || vg.density_layers_hash != fd.density_layers_hash) {
            diag = "MissingRequiredResource phase=grain_payload field=density_layers";
            return hipSuccess;
        }
        if (n <= 0 || n > kGrainCurveMax) {
            diag = "UnsupportedInput grain sub-layer curves longer than 256 samples";
            return hipSuccess;
        }
    } else if (vg.density_layers_hash != 0) {
        diag = "ResourceDescriptorMismatch phase=grain_payload field=inactive_density_layers";
        return hipSuccess;
    }
    // juicer_cuda_apply_visual_grain preconditions: a refused launch.
    auto reject = [&]() {
        launch_rejected = true;
        diag = "print visual grain launch failed";
        return hipSuccess;
    };
    for (int c = 0; c < 3; ++c) {
        if (!std::isfinite(vg.particle_density_min_cmy[c]) || !finite_positive(vg.density_max_cmy[c]) ||
            !finite_positive(vg.n_particles_cmy[c]) || !finite_positive(vg.od_particle_cmy[c]))
            return reject();
        if (!sub) continue;
        for (int l = 0; l < 3; ++l)
            if (!std::isfinite(vg.density_min_layers[l][c]) || !finite_positive(vg.density_max_layers[l][c]) ||
                !finite_positive(vg.n_particles_layers[l][c]) || !finite_positive(vg.od_particle_layers[l][c]))
                return reject();
    }
    const float wf = vg.size_mix_weights[0], wm = vg.size_mix_weights[1], wc = vg.size_mix_weights[2];
    if (!std::isfinite(wf) || !std::isfinite(wm) || !std::isfinite(wc) || wf < 0.0f || wm < 0.0f || wc < 0.0f ||
        std::fabs((wf + wm + wc) - 1.0f) > 1e-5f || !std::isfinite(vg.size_mix_scale) || !(vg.size_mix_scale >= 1.0f) ||
        !finite_positive(vg.size_mix_gain) || !std::isfinite(vg.amplitude) || vg.amplitude < 0.0f)
        return reject();
    if ((wm > 0.0f && vg.correlation[1].radius <= 0) || (wc > 0.0f && vg.correlation[2].radius <= 0)) return reject();
    if (!std::isfinite(vg.chroma_mix) || vg.chroma_mix < 0.0f || vg.chroma_mix > 1.0f ||
        !std::isfinite(vg.chroma_shared_weight) || vg.chroma_shared_weight < 0.0f ||
        !std::isfinite(vg.chroma_independent_weight) || vg.chroma_independent_weight < 0.0f)
        return reject();

    GrainDev& g = out.base;
    const void* ptr = nullptr;
    // static noise, uploaded once per owner
    const size_t stbn_bytes = static_cast<size_t>(nz.stbn_width) * nz.stbn_height * nz.stbn_frames;
    hipError_t e = dev.table(span_identity(nz.stbn, 0x57B4ull), nz.stbn.data, stbn_bytes, &ptr);
    if (e != hipSuccess) return e;
    g.stbn = static_cast<const uint8_t*>(ptr);
    const size_t wang_bytes = static_cast<size_t>(nz.wang_width) * nz.wang_height * nz.wang_tile_count;
    e = dev.table(span_identity(nz.wang_tiles, 0x3A46ull), nz.wang_tiles.data, wang_bytes, &ptr);
    if (e != hipSuccess) return e;
    g.wang_tiles = static_cast<const uint8_t*>(ptr);
    const size_t lut_bytes = static_cast<size_t>(nz.wang_colors) * nz.wang_colors * nz.wang_colors * nz.wang_colors;
    e = dev.table(hash_bytes(nz.wang_lut.data, lut_bytes, 0x1077ull), nz.wang_lut.data, lut_bytes, &ptr);
    if (e != hipSuccess) return e;
    g.wang_lut = static_cast<const uint8_t*>(ptr);
    g.stbn_w = nz.stbn_width;
    g.stbn_h = nz.stbn_height;
    g.stbn_frames = nz.stbn_frames;
    g.stbn_mask = (g.stbn_w & (g.stbn_w - 1)) == 0 && (g.stbn_h & (g.stbn_h - 1)) == 0;
    g.stbn_off_x = static_cast<int>(fnv_u64s({vg.session_seed, 0xA5u}) % static_cast<uint64_t>(g.stbn_w));
    g.stbn_off_y = static_cast<int>(fnv_u64s({vg.session_seed, 0x5Au}) % static_cast<uint64_t>(g.stbn_h));
    {
        const int64_t phase = static_cast<int64_t>(vg.session_seed % static_cast<uint64_t>(g.stbn_frames));
        int f = static_cast<int>((vg.frame0 + phase) % g.stbn_frames);
        if (f < 0) f += g.stbn_frames;
        g.stbn_frame = f;
    }
    g.frame_index = vg.frame0;
    g.wang_w = nz.wang_width;
    g.wang_h = nz.wang_height;
    g.wang_count = static_cast<int>(nz.wang_tile_count);
    g.wang_colors = nz.wang_colors;
    g.wang_cell_mm = vg.wang_cell_mm;
    g.pixel_size_um = vg.pixel_size_um;
    g.session_seed = vg.session_seed;
    g.clip_token = vg.clip_token;
    g.origin_x = vg.render_extent.x;
    g.origin_y = vg.render_extent.y;
    g.pitch_px = vg.pitch_px;
    g.time_alpha = vg.frame_alpha;
    g.seed_base = vg.seed_base;
    g.seed_base_next = vg.seed_base_next;
    g.n_sublayers = static_cast<int>(vg.sublayer_count);
    g.positive_film = p.capture_polarity == FJ_POLARITY_POSITIVE;
    g.curve_n = n;
    for (int c = 0; c < 3; ++c) {
        g.density_min[c] = vg.particle_density_min_cmy[c];
        g.uniformity[c] = vg.uniformity_cmy[c];
        g.density_max[c] = vg.density_max_cmy[c];
        g.n_particles[c] = vg.n_particles_cmy[c];
        g.od_particle[c] = vg.od_particle_cmy[c];
        g.axis_finite[c] = sub && (vg.flags & (1u << (3u + c))) != 0;
        for (int b = 0; b < 16; ++b) g.axis_prefix[c][b] = sub ? vg.layer_axis_block_prefix_max[c][b] : 0.0f;
        for (int l = 0; l < 3; ++l) {
            g.dmin_l[l][c] = vg.density_min_layers[l][c];
            g.dmax_l[l][c] = vg.density_max_layers[l][c];
            g.n_l[l][c] = vg.n_particles_layers[l][c];
            g.od_l[l][c] = vg.od_particle_layers[l][c];
        }
    }
    if (sub) {
        // Interpolation axis: the development's density curve of each dye
        // (the densR/G/B base curves, ProcessRoot.cpp:3537-3561).
        std::vector<float> axes(static_cast<size_t>(3) * n);
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < n; ++i) axes[static_cast<size_t>(c) * n + i] = fd.density_rgb.data[i * 3 + c];
        e = dev.table(hash_bytes(axes.data(), axes.size() * 4, 0xA715ull), axes.data(), axes.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        for (int c = 0; c < 3; ++c) g.base_curve[c] = static_cast<const float*>(ptr) + static_cast<size_t>(c) * n;
        std::vector<float> layers(static_cast<size_t>(9) * n);
        for (int l = 0; l < 3; ++l)
            for (int c = 0; c < 3; ++c)
                std::memcpy(&layers[static_cast<size_t>(l * 3 + c) * n], fd.density_layers[l][c].data, sizeof(float) * n);
        e = dev.table(hash_bytes(layers.data(), layers.size() * 4, 0x1A7Eull), layers.data(), layers.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        for (int l = 0; l < 3; ++l)
            for (int c = 0; c < 3; ++c)
                g.layer_curve[l][c] = static_cast<const float*>(ptr) + static_cast<size_t>(l * 3 + c) * n;
    }
    grain_frame_uniforms(vg, g);
    auto kernel = [&](const FjGrainGaussian& gg, GrainKernel& k) -> hipError_t {
        k = GrainKernel{};
        if (gg.radius <= 0) return hipSuccess;
        std::vector<float> w = kernel_weights(gg.radius, gg.sigma_px);
        if (w.empty()) return hipErrorInvalidValue;
        const void* wp = nullptr;
        const hipError_t r = dev.table(hash_bytes(w.data(), w.size() * 4, 0x6A55ull), w.data(), w.size() * 4, &wp);
        if (r != hipSuccess) return r;
        k.weights = static_cast<const float*>(wp);
        k.radius = gg.radius;
        k.sigma = gg.sigma_px;
        return hipSuccess;
    };
    for (int k = 0; k < 3; ++k) {
        e = kernel(vg.correlation[k], out.corr[k]);
        if (e != hipSuccess) return e;
    }
    for (int l = 0; l < 3; ++l)
        for (int c = 0; c < 3; ++c) {
            e = kernel(vg.dye_cloud[l][c], out.dye[l][c]);
            if (e != hipSuccess) return e;
        }
    out.w_fine = wf;
    out.w_mid = wm;
    out.w_coarse = wc;
    out.size_mix_scale = vg.size_mix_scale;
    out.size_mix_gain = vg.size_mix_gain;
    out.amplitude = vg.amplitude;
    out.chroma_mix = vg.chroma_mix;
    out.shared_w = vg.chroma_shared_weight;
    out.ind_w = vg.chroma_independent_weight;
    out.debug_scale = vg.debug_scale;
    out.debug_view = static_cast<int>(vg.debug_view);
    out.sublayers = sub;
    ok = true;
    return hipSuccess;
}

namespace {

struct Work {
    float *tmp, *scale, *delta, *layer, *shared;
    long long n;
    int width, height;
};

hipError_t elt(int op, float* a, const float* b, long long n, hipStream_t s, int init = 0, float w = 0.0f,
               float bias = 0.0f, float sc = 0.0f) {
    GrainEltArgs e;
    e.op = op;
    e.a = a;
    e.b = b;
    e.n = n;
    e.init = init;
    e.w = w;
    e.bias = bias;
    e.s = sc;
    return launch_grain_elt(e, s);
}

hipError_t blur_pair(const float* in, float* tmp, float* out, const GrainKernel& k, const Work& wk, int weighted,
                     float weight, int init, const float* sub, float bias, hipStream_t s, float gain = 1.0f) {
    GrainBlurArgs h;
    h.in = in;
    h.out = tmp;
    h.weights = k.weights;
    h.radius = k.radius;
    h.width = wk.width;
    h.height = wk.height;
    hipError_t e = launch_grain_blur_h(h, s);
    if (e != hipSuccess) return e;
    GrainBlurArgs v = h;
    v.in = tmp;
    v.out = out;
    v.weighted = weighted;
    v.weight = weight;
    v.initialize = init;
    v.sub = sub;
    v.bias = bias;
    v.gain = gain;
    return launch_grain_blur_v(v, s);
}

// One grain scale accumulated (weighted) into `out`; when the correlation
// blur runs, its epilogue also multiplies by `gain` (1: none).
hipError_t generate_scale(const GrainPlan& plan, const GrainDev& base, const float* dens, int c, uint64_t salt,
                          float factor, const GrainKernel& corr, bool apply_corr, float weight, bool init,
                          const Work& wk, float* out, hipStream_t s, float gain = 1.0f) {
    if (!std::isfinite(factor) || !(factor > 0.0f)) return hipErrorInvalidValue;
    GrainSampleArgs a{};
    a.g = base;
    GrainDev& g = a.g;
    g.seed_base ^= salt;
    g.seed_base_next ^= salt;
    for (int k = 0; k < 3; ++k) {
        g.n_particles[k] = base.n_particles[k] / factor;
        g.od_particle[k] = base.od_particle[k] * factor;
        for (int l = 0; l < 3; ++l) {
            g.n_l[l][k] = base.n_l[l][k] / factor;
            g.od_l[l][k] = base.od_l[l][k] * factor;
        }
    }
    a.density = dens;
    a.channel = c;
    a.width = wk.width;
    a.height = wk.height;
    const float bias = -base.density_min[c];
    a.bias = bias;
    hipError_t e;
    bool formed = false;
    if (plan.sublayers) {
        const GrainKernel& d0 = plan.dye[0][c];
        const GrainKernel& d1 = plan.dye[1][c];
        const GrainKernel& d2 = plan.dye[2][c];
        const bool same = d0.radius == d1.radius && d0.radius == d2.radius &&
                          (d0.radius == 0 || (d0.sigma == d1.sigma && d0.sigma == d2.sigma));
        if (same) {
            if (d0.radius > 0) {
                a.out = wk.layer;
                e = launch_grain_samples(a, kGrainTriplet, s);
                if (e != hipSuccess) return e;
                e = blur_pair(wk.layer, wk.tmp, wk.scale, d0, wk, 0, 0.0f, 1, dens, bias, s);
            } else {
                a.out = wk.scale;
                a.form_delta = 1;
                e = launch_grain_samples(a, kGrainTriplet, s);
            }
            if (e != hipSuccess) return e;
            formed = true;
        } else {
            for (int l = 0; l < 3; ++l) {
                a.layer = l;
                a.out = wk.layer;
                e = launch_grain_samples(a, kGrainLayer, s);
                if (e != hipSuccess) return e;
                const GrainKernel& d = plan.dye[l][c];
                if (d.radius > 0) {
                    e = blur_pair(wk.layer, wk.tmp, wk.scale, d, wk, 0, 0.0f, l == 0, l == 2 ? dens : nullptr,
                                  l == 2 ? bias : 0.0f, s);
                    formed = l == 2;
                } else {
                    e = elt(kEltAccumulate, wk.scale, wk.layer, wk.n, s, l == 0);
                }
                if (e != hipSuccess) return e;
            }
        }
    } else {
        a.out = wk.scale;
        a.form_delta = 1;
        e = launch_grain_samples(a, kGrainSimple, s);
        if (e != hipSuccess) return e;
        formed = true;
    }
    if (!formed) {
        e = bias != 0.0f ? elt(kEltFormDelta, wk.scale, dens, wk.n, s, 0, 0.0f, bias)
                         : elt(kEltSubtract, wk.scale, dens, wk.n, s);
        if (e != hipSuccess) return e;
    }
    if (apply_corr && corr.radius > 0)
        return blur_pair(wk.scale, wk.tmp, out, corr, wk, 1, weight, init ? 1 : 0, nullptr, 0.0f, s, gain);
    return elt(kEltAccWeighted, out, wk.scale, wk.n, s, init ? 1 : 0, weight);
}

// The fine / mid / coarse scales summed into `out`, times the size-mix gain.
// The gain rides on the last scale's correlation blur when that blur runs
// (san(v * gain) there is the separate scale pass, bit for bit).
hipError_t generate_mix(const GrainPlan& plan, const GrainDev& base, const float* dens, int c, const Work& wk,
                        float* out, hipStream_t s) {
    struct Scale {
        bool on;
        uint64_t salt;
        float factor;
        const GrainKernel& corr;
        bool apply;
        float weight;
    };
    const Scale scales[3] = {
        {plan.w_fine > 0.0f, kSaltFine, 1.0f, plan.corr[0], plan.corr[0].radius > 0, plan.w_fine},
        {plan.w_mid > 0.0f, kSaltMid, std::sqrt(plan.size_mix_scale), plan.corr[1], true, plan.w_mid},
        {plan.w_coarse > 0.0f, kSaltCoarse, plan.size_mix_scale, plan.corr[2], true, plan.w_coarse},
    };
    int last = -1;
    for (int i = 0; i < 3; ++i)
        if (scales[i].on) last = i;
    const float gain = std::isfinite(plan.size_mix_gain) ? plan.size_mix_gain : 0.0f;
    bool init = true, gained = false;
    for (int i = 0; i < 3; ++i) {
        const Scale& sc = scales[i];
        if (!sc.on) continue;
        const bool fuse = i == last && sc.apply && sc.corr.radius > 0;
        hipError_t e = generate_scale(plan, base, dens, c, sc.salt, sc.factor, sc.corr, sc.apply, sc.weight, init, wk,
                                      out, s, fuse ? gain : 1.0f);
        if (e != hipSuccess) return e;
        init = false;
        gained = gained || fuse;
    }
    if (gained) return hipSuccess;
    return elt(kEltScale, out, nullptr, wk.n, s, 0, 0.0f, 0.0f, plan.size_mix_gain);
}

}  // namespace

hipError_t run_grain(const GrainPlan& plan, float* const dens[3], int width, int height, float* scratch,
                     hipStream_t stream) {
    Work wk;
    wk.n = static_cast<long long>(width) * height;
    wk.width = width;
    wk.height = height;
    wk.tmp = scratch;
    wk.scale = scratch + wk.n;
    wk.delta = scratch + 2 * wk.n;
    wk.layer = scratch + 3 * wk.n;
    wk.shared = scratch + 4 * wk.n;
    hipError_t e;
    GrainAvgArgs avg{};
    for (int c = 0; c < 3; ++c) avg.d[c] = dens[c];
    avg.n = wk.n;
    if (plan.debug_view == FJ_GRAIN_DEBUG_MEAN_DENSITY) {
        avg.offset = 0.0f;
        avg.scale = plan.debug_scale * 4.0f;
        return launch_grain_avg3(avg, stream);
    }
    if (plan.debug_view >= 2 && plan.debug_view <= 5) {
        const bool coarse = plan.debug_view == 3 || plan.debug_view == 5;
        const bool raw = plan.debug_view == 4 || plan.debug_view == 5;
        for (int c = 0; c < 3; ++c) {
            if (!coarse || plan.w_coarse > 0.0f) {
                const GrainKernel& k = coarse ? plan.corr[2] : plan.corr[0];
                e = generate_scale(plan, plan.base, dens[c], c, coarse ? kSaltCoarse : kSaltFine,
                                   coarse ? plan.size_mix_scale : 1.0f, k, !raw && k.radius > 0, 1.0f, true, wk,
                                   wk.delta, stream);
            } else {
                e = elt(kEltClear, wk.delta, nullptr, wk.n, stream);
            }
            if (e != hipSuccess) return e;
            e = elt(kEltReconstruct, dens[c], wk.delta, wk.n, stream, 0, 0.0f, 0.0f, plan.amplitude);
            if (e != hipSuccess) return e;
        }
        avg.offset = 0.5f;
        avg.scale = plan.debug_scale;
        return launch_grain_avg3(avg, stream);
    }
    const bool shared = (plan.debug_view == 0 || plan.debug_view == 1) && plan.chroma_mix < 0.999f &&
                        plan.shared_w > 0.0f;
    if (shared) {
        GrainDev sg = plan.base;
        sg.seed_base ^= kSaltShared;
        sg.seed_base_next ^= kSaltShared;
        e = generate_mix(plan, sg, dens[1], 1, wk, wk.shared, stream);
        if (e != hipSuccess) return e;
    }
    for (int c = 0; c < 3; ++c) {
        const bool independent = plan.ind_w > 0.0f || !shared;
        if (independent) {
            e = generate_mix(plan, plan.base, dens[c], c, wk, wk.delta, stream);
            if (e != hipSuccess) return e;
        }
        // With the shared part: the mix and the reconstruction in one pass
        // (the independent delta is read only when it is used).
        GrainEltArgs r;
        r.a = dens[c];
        r.b = wk.delta;
        r.n = wk.n;
        r.add_mean = plan.debug_view == 0;
        if (shared) {
            r.op = kEltMixReconstruct;
            r.shared = wk.shared;
            r.use_ind = plan.ind_w > 0.0f;
            r.sw = plan.shared_w;
            r.iw = plan.ind_w;
            r.amp = plan.amplitude;
            r.s = 1.0f;
        } else {
            r.op = kEltReconstruct;
            r.s = plan.amplitude;
        }
        e = launch_grain_elt(r, stream);
        if (e != hipSuccess) return e;
    }
    if (plan.debug_view != 0) {
        avg.offset = 0.5f;
        avg.scale = plan.debug_scale;
        return launch_grain_avg3(avg, stream);
    }
    return hipSuccess;
}

}  // namespace fjx
