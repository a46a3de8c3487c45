// fj_frame_descriptors.cpp — per-frame descriptor composition for the call
// scope. The descriptors are identities plus validated copies of the host's
// frame math; their hashes are FNV-1a over the same fields, in the same
// order and widths, as the reference's descriptor structs (so a host that
// caches by descriptor hash sees the same identities).
#include "fj_frame_descriptors.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace fjx {

namespace {

constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ull;
constexpr uint64_t kFnvPrime = 0x100000001b3ull;

// FNV-1a accumulator over the little-endian bytes of fixed-width values.
struct Fnv {
    uint64_t h = kFnvOffset;
    void bytes(const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= kFnvPrime;
        }
    }
    template <typename T>
    void put(T v) {
        bytes(&v, sizeof(v));
    }
    void extent(const FjExtent& e) {
        put<int32_t>(e.x);
        put<int32_t>(e.y);
        put<int32_t>(e.width);
        put<int32_t>(e.height);
    }
};

bool zero_bits(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u == 0;
}
bool zero_bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return u == 0;
}
bool zero_extent(const FjExtent& e) { return e.x == 0 && e.y == 0 && e.width == 0 && e.height == 0; }
bool same_extent(const FjExtent& a, const FjExtent& b) {
    return a.x == b.x && a.y == b.y && a.width == b.width && a.height == b.height;
}

// --- spatial DIR ------------------------------------------------------------

// Scratch tiers and operators as the reference's one-byte enums
// (src/RenderValues.h:108-128).
enum : uint8_t { kTier1F = 1, kTier1IChannels = 2, kTier2 = 3 };
enum : uint8_t { kSupported = 0 };

bool dir_math_empty(const FjDirFrameMath& m) {
    if (m.present != 0 || m.component_count != 0 || m.recipe_hash != 0 || !zero_extent(m.render_extent) ||
        !zero_extent(m.full_frame_extent) || !zero_extent(m.filter_domain_extent))
        return false;
    for (const FjDirGaussian& g : m.components) {
        if (!zero_bits(g.sigma_px) || !zero_bits(g.weight) || g.radius != 0 || g.reference_operator != 0 ||
            !zero_bits(g.feedforward) || !zero_bits(g.normalization_denominator) ||
            !zero_bits(g.feedforward_numerator) || !zero_bits(g.boundary_truncation_accuracy) ||
            !zero_bits(g.boundary_certification_tolerance) || g.boundary_derivation_version != 0)
            return false;
        for (int k = 0; k < 3; ++k)
            if (!zero_bits(g.feedback[k]) || !zero_bits(g.feedback_numerators[k])) return false;
    }
    return true;
}

// --- visual grain -----------------------------------------------------------

constexpr uint32_t kGrainAllowedFlags =
    FJ_GRAIN_ACTIVE | FJ_GRAIN_SUBLAYERS | FJ_GRAIN_AXIS_FINITE_C | FJ_GRAIN_AXIS_FINITE_M | FJ_GRAIN_AXIS_FINITE_Y;

bool gaussian_empty(const FjGrainGaussian& g) { return zero_bits(g.sigma_px) && g.radius == 0 && g.hash == 0; }

// A Gaussian of the frame math: absent (all zero), or radius 1..75 with a
// finite positive sigma.
bool gaussian_ok(const FjGrainGaussian& g) {
    if (g.radius == 0) return zero_bits(g.sigma_px) && g.hash == 0;
    return g.radius >= 1 && g.radius <= 75 && std::isfinite(g.sigma_px) && g.sigma_px > 0.0f;
}

bool grain_math_empty(const FjGrainFrameMath& m) {
    if (m.present != 0 || m.capture_polarity != 0 || !zero_extent(m.render_extent) ||
        !zero_extent(m.full_frame_extent) || !zero_bits(m.pixel_size_um) || m.frame0 != 0 ||
        !zero_bits(m.frame_alpha) || m.seed_base != 0 || m.seed_base_next != 0 || m.session_seed != 0 ||
        m.clip_token != 0 || m.pitch_px != 0 || m.breathing_period_frames != 0 || m.clump_morph_period_frames != 0 ||
        !zero_bits(m.wang_cell_mm) || !zero_bits(m.breathing_amplitude) || !zero_bits(m.breathing_cell_um_small) ||
        !zero_bits(m.breathing_cell_um_large) || !zero_bits(m.breathing_mix) ||
        !zero_bits(m.breathing_drift_um_per_frame) || !zero_bits(m.debug_scale) || m.requires_full_frame != 0 ||
        m.recipe_hash != 0 || m.density_layers_hash != 0 || !zero_bits(m.size_mix_gain))
        return false;
    for (int c = 0; c < 3; ++c) {
        if (!zero_bits(m.density_max_cmy[c]) || !zero_bits(m.n_particles_cmy[c]) || !zero_bits(m.od_particle_cmy[c]) ||
            !zero_bits(m.size_mix_weights[c]) || !gaussian_empty(m.correlation[c]))
            return false;
        for (int l = 0; l < 3; ++l)
            if (!zero_bits(m.density_min_layers[c][l]) || !zero_bits(m.density_max_layers[c][l]) ||
                !zero_bits(m.n_particles_layers[c][l]) || !zero_bits(m.od_particle_layers[c][l]) ||
                !gaussian_empty(m.dye_cloud[c][l]))
                return false;
    }
    return true;
}

// Scratch shape of the grain schedule (visual_grain_scratch_shape,
// src/FilmEffectsFrameDescriptors.cpp:284-292; enum :14-20): 1 streamed,
// 2 streamed + shared chroma, 3 sub-layers, 4 sub-layers + shared chroma.
uint8_t grain_scratch_shape(const FjGrainCoefficients& k) {
    if ((k.flags & FJ_GRAIN_ACTIVE) == 0) return 0;
    const bool shared = (k.debug_view == FJ_GRAIN_DEBUG_OFF || k.debug_view == FJ_GRAIN_DEBUG_DELTA_MIX) &&
                        k.chroma_mix < 0.999f && k.chroma_shared_weight > 0.0f;
    if ((k.flags & FJ_GRAIN_SUBLAYERS) != 0) return shared ? 4 : 3;
    return shared ? 2 : 1;
}

uint64_t grain_hash(const FjGrainFrameMath& m, uint8_t polarity, uint8_t shape, bool full_frame) {
    Fnv f;
    f.put<uint8_t>(1);  // active
    f.put<uint8_t>(polarity);
    f.extent(m.render_extent);
    f.extent(m.full_frame_extent);
    f.put<float>(m.pixel_size_um);
    f.put<int64_t>(m.frame0);
    f.put<float>(m.frame_alpha);
    f.put<uint64_t>(m.seed_base);
    f.put<uint64_t>(m.seed_base_next);
    f.put<uint64_t>(m.session_seed);
    f.put<uint64_t>(m.clip_token);
    f.put<int32_t>(m.pitch_px);
    f.put<int32_t>(m.breathing_period_frames);
    f.put<int32_t>(m.clump_morph_period_frames);
    f.put<float>(m.wang_cell_mm);
    f.put<float>(m.breathing_amplitude);
    f.put<float>(m.breathing_cell_um_small);
    f.put<float>(m.breathing_cell_um_large);
    f.put<float>(m.breathing_mix);
    f.put<float>(m.breathing_drift_um_per_frame);
    f.put<float>(m.debug_scale);
    f.bytes(m.density_max_cmy, sizeof(m.density_max_cmy));
    f.bytes(m.n_particles_cmy, sizeof(m.n_particles_cmy));
    f.bytes(m.od_particle_cmy, sizeof(m.od_particle_cmy));
    f.bytes(m.density_min_layers, sizeof(m.density_min_layers));
    f.bytes(m.density_max_layers, sizeof(m.density_max_layers));
    f.bytes(m.n_particles_layers, sizeof(m.n_particles_layers));
    f.bytes(m.od_particle_layers, sizeof(m.od_particle_layers));
    for (const FjGrainGaussian& g : m.correlation) f.put<uint64_t>(g.hash);
    for (const auto& row : m.dye_cloud)
        for (const FjGrainGaussian& g : row) f.put<uint64_t>(g.hash);
    for (int i = 0; i < 3; ++i) f.put<float>(m.size_mix_weights[i]);
    f.put<float>(m.size_mix_gain);
    f.put<uint8_t>(shape);
    f.put<uint8_t>(full_frame ? 1 : 0);
    f.put<uint64_t>(m.recipe_hash);
    f.put<uint64_t>(m.density_layers_hash);
    return f.h;
}

}  // namespace

bool compose_spatial_dir(const FjDirFrameMath& math, float pixel_size_um, FjSpatialDir& out) {
    out = FjSpatialDir{};
    if (math.present == 0) return dir_math_empty(math);
    if (math.present != 1 || math.component_count == 0 || math.component_count > 4) return false;
    const int count = static_cast<int>(math.component_count);
    // Scratch tiers: every weighted component targets the cached-log-raw
    // tier; the build tier is per-channel when a weighted component is a
    // recursive (YVV) one.
    bool uses_yvv = false;
    uint8_t comp_tier[4] = {};
    for (int i = 0; i < count; ++i) {
        const FjDirGaussian& g = math.components[i];
        const bool yvv = static_cast<uint8_t>(g.reference_operator) == FJ_DIR_OPERATOR_YVV_REFLECT;
        comp_tier[i] = g.weight > 0.0f ? kTier2 : (yvv ? kTier1IChannels : kTier1F);
        if (g.weight > 0.0f && yvv) uses_yvv = true;
    }
    const uint8_t build_tier = uses_yvv ? kTier1IChannels : kTier1F;
    const int32_t roles[4] = {3, 3, uses_yvv ? 3 : 1, 0};         // raw, filtered, temp, cached log-raw
    const int32_t target_roles[4] = {3, 3, uses_yvv ? 3 : 1, 3};
    Fnv f;
    f.put<uint64_t>(math.recipe_hash);
    f.put<float>(pixel_size_um);
    f.put<uint8_t>(build_tier);
    f.put<uint8_t>(kTier2);
    f.put<uint8_t>(kSupported);
    f.extent(math.render_extent);
    f.extent(math.full_frame_extent);
    f.extent(math.filter_domain_extent);
    f.put<int32_t>(count);
    for (int i = 0; i < count; ++i) {
        const FjDirGaussian& g = math.components[i];
        f.put<float>(g.sigma_px);
        f.put<float>(g.weight);
        f.put<int32_t>(g.radius);
        f.put<uint8_t>(static_cast<uint8_t>(g.reference_operator));
        f.put<uint8_t>(comp_tier[i]);
        if (static_cast<uint8_t>(g.reference_operator) == FJ_DIR_OPERATOR_YVV_REFLECT) {
            f.put<double>(g.feedforward);
            for (double v : g.feedback) f.put<double>(v);
            f.put<double>(g.normalization_denominator);
            f.put<double>(g.feedforward_numerator);
            for (double v : g.feedback_numerators) f.put<double>(v);
            f.put<double>(g.boundary_truncation_accuracy);
            f.put<double>(g.boundary_certification_tolerance);
            f.put<uint32_t>(g.boundary_derivation_version);
        }
    }
    for (int32_t v : roles) f.put<int32_t>(v);
    for (int32_t v : target_roles) f.put<int32_t>(v);
    if (f.h == 0) return false;
    out.render_extent = math.render_extent;
    out.full_frame_extent = math.full_frame_extent;
    out.filter_domain_extent = math.filter_domain_extent;
    out.component_count = static_cast<size_t>(count);
    for (int i = 0; i < count; ++i) out.components[i] = math.components[i];
    out.recipe_hash = math.recipe_hash;
    out.hash = f.h;
    return true;
}

bool compose_visual_grain(const FjGrainCoefficients& k, uint32_t capture_polarity, const FjGrainFrameMath& m,
                          FjVisualGrain& out, std::string& diagnostic) {
    out = FjVisualGrain{};
    diagnostic.clear();
    const char* kCoefficients = "ResourceDescriptorMismatch phase=grain_descriptor field=coefficients";
    if ((k.flags & ~kGrainAllowedFlags) != 0 || k.debug_view > FJ_GRAIN_DEBUG_MEAN_DENSITY ||
        k.sublayer_count > static_cast<size_t>(INT_MAX)) {
        diagnostic = kCoefficients;
        return false;
    }
    if ((k.flags & FJ_GRAIN_ACTIVE) == 0) {
        if (k.flags != 0 || k.debug_view != 0 || k.sublayer_count != 0 || !grain_math_empty(m)) {
            diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=inactive";
            return false;
        }
        return true;
    }
    if (capture_polarity > 1 || k.sublayer_count == 0 ||
        (((k.flags & FJ_GRAIN_SUBLAYERS) != 0) != (m.density_layers_hash != 0))) {
        diagnostic = kCoefficients;
        return false;
    }
    if (m.present != 1 || m.capture_polarity > 1 || m.capture_polarity != capture_polarity || m.recipe_hash == 0 ||
        m.requires_full_frame > 1 || m.render_extent.width <= 0 || m.render_extent.height <= 0 ||
        m.full_frame_extent.width <= 0 || m.full_frame_extent.height <= 0 || m.pitch_px != m.full_frame_extent.height) {
        diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=rust_result";
        return false;
    }
    bool full_frame = false;
    for (const FjGrainGaussian& g : m.correlation) {
        if (!gaussian_ok(g)) {
            diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=gaussian";
            return false;
        }
        full_frame = full_frame || g.radius > 0;
    }
    for (const auto& row : m.dye_cloud)
        for (const FjGrainGaussian& g : row) {
            if (!gaussian_ok(g)) {
                diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=gaussian";
                return false;
            }
            full_frame = full_frame || g.radius > 0;
        }
    if (m.requires_full_frame != static_cast<uint32_t>(full_frame) ||
        (full_frame && !same_extent(m.render_extent, m.full_frame_extent))) {
        diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=full_frame";
        return false;
    }
    const uint64_t hash = grain_hash(m, static_cast<uint8_t>(capture_polarity), grain_scratch_shape(k), full_frame);
    if (hash == 0) {
        diagnostic = "ResourceDescriptorMismatch phase=grain_descriptor field=hash";
        return false;
    }
    out.flags = k.flags | (full_frame ? FJ_GRAIN_FULL_FRAME : 0u);
    out.debug_view = k.debug_view;
    out.render_extent = m.render_extent;
    out.full_frame_extent = m.full_frame_extent;
    out.pixel_size_um = m.pixel_size_um;
    out.frame0 = m.frame0;
    out.frame_alpha = m.frame_alpha;
    out.seed_base = m.seed_base;
    out.seed_base_next = m.seed_base_next;
    out.session_seed = m.session_seed;
    out.clip_token = m.clip_token;
    out.pitch_px = m.pitch_px;
    out.breathing_period_frames = m.breathing_period_frames;
    out.clump_morph_period_frames = m.clump_morph_period_frames;
    out.wang_cell_mm = m.wang_cell_mm;
    out.breathing_amplitude = m.breathing_amplitude;
    out.breathing_cell_um_small = m.breathing_cell_um_small;
    out.breathing_cell_um_large = m.breathing_cell_um_large;
    out.breathing_mix = m.breathing_mix;
    out.breathing_drift_um_per_frame = m.breathing_drift_um_per_frame;
    out.debug_scale = m.debug_scale;
    for (int c = 0; c < 3; ++c) {
        out.density_max_cmy[c] = m.density_max_cmy[c];
        out.n_particles_cmy[c] = m.n_particles_cmy[c];
        out.od_particle_cmy[c] = m.od_particle_cmy[c];
        out.size_mix_weights[c] = m.size_mix_weights[c];
        out.particle_density_min_cmy[c] = k.particle_density_min_cmy[c];
        out.uniformity_cmy[c] = k.uniformity_cmy[c];
        out.correlation[c] = m.correlation[c];
        for (int b = 0; b < 16; ++b) out.layer_axis_block_prefix_max[c][b] = k.layer_axis_block_prefix_max[c][b];
        for (int l = 0; l < 3; ++l) {
            out.density_min_layers[c][l] = m.density_min_layers[c][l];
            out.density_max_layers[c][l] = m.density_max_layers[c][l];
            out.n_particles_layers[c][l] = m.n_particles_layers[c][l];
            out.od_particle_layers[c][l] = m.od_particle_layers[c][l];
            out.dye_cloud[c][l] = m.dye_cloud[c][l];
        }
    }
    out.size_mix_gain = m.size_mix_gain;
    out.sublayer_count = k.sublayer_count;
    out.micro_structure[0] = k.micro_structure[0];
    out.micro_structure[1] = k.micro_structure[1];
    out.clump_temporal_mix = k.clump_temporal_mix;
    out.size_mix_scale = k.size_mix_scale;
    out.amplitude = k.amplitude;
    out.chroma_mix = k.chroma_mix;
    out.chroma_shared_weight = k.chroma_shared_weight;
    out.chroma_independent_weight = k.chroma_independent_weight;
    out.recipe_hash = m.recipe_hash;
    out.density_layers_hash = m.density_layers_hash;
    out.hash = hash;
    return true;
}

bool effects_math_empty(const FjEffectsFrameMath& m) {
    if (m.present != 0 || m.flags != 0 || !zero_extent(m.render_extent) || !zero_extent(m.full_frame_extent) ||
        !zero_bits(m.sample_step_x_mm) || !zero_bits(m.sample_step_y_mm) || m.roi_offset_x != 0 ||
        m.roi_offset_y != 0 || m.gate_width != 0 || m.gate_height != 0 || m.session_seed != 0 || m.clip_token != 0 ||
        !zero_bits(m.weave_dx_px) || !zero_bits(m.weave_dy_px) || !zero_bits(m.weave_cos_rot) ||
        !zero_bits(m.weave_sin_rot) || m.recipe_hash != 0 || m.hash != 0)
        return false;
    for (const FjDefectOrigin& o : m.origins)
        if (o.cell_x != 0 || o.cell_y != 0 || !zero_bits(o.local_x_mm) || !zero_bits(o.local_y_mm)) return false;
    return true;
}

uint32_t compose_frame_descriptors(uint32_t route, const FjFrame& f, const FjFrameMathInput& in,
                                   FjNativeFrameDescriptors& out, std::string& diagnostic) {
    out = FjNativeFrameDescriptors{};
    diagnostic.clear();
    const char* kMismatch = "ResourceDescriptorMismatch phase=frame_descriptors field=rust_result";
    if (route > FJ_ROUTE_POSITIVE_PRINT) {
        diagnostic = "UnsupportedInput route";
        return FJ_STATUS_UNSUPPORTED_INPUT;
    }
    const FjExtent render{f.render_window.x1, f.render_window.y1, f.render_window.x2 - f.render_window.x1,
                          f.render_window.y2 - f.render_window.y1};
    const FjExtent full{f.full_frame_extent.x1, f.full_frame_extent.y1, f.full_frame_extent.x2 - f.full_frame_extent.x1,
                        f.full_frame_extent.y2 - f.full_frame_extent.y1};
    uint32_t grain_px, frame_px;
    std::memcpy(&grain_px, &in.grain.pixel_size_um, 4);
    std::memcpy(&frame_px, &f.pixel_size_um, 4);
    if ((in.dir.present != 0 &&
         (!same_extent(in.dir.render_extent, render) || !same_extent(in.dir.full_frame_extent, full))) ||
        (in.grain.present != 0 && (!same_extent(in.grain.render_extent, render) ||
                                   !same_extent(in.grain.full_frame_extent, full) || grain_px != frame_px)) ||
        (in.effects.present != 0 && (!same_extent(in.effects.render_extent, render) ||
                                     !same_extent(in.effects.full_frame_extent, f.effects_geometry.pixel_definition)))) {
        diagnostic = kMismatch;
        return FJ_STATUS_PREPARATION_FAILURE;
    }
    FjNativeFrameDescriptors d{};
    if (!compose_spatial_dir(in.dir, f.pixel_size_um, d.spatial_dir)) {
        diagnostic = "ResourceDescriptorMismatch phase=dir_descriptor field=rust_result";
        return FJ_STATUS_PREPARATION_FAILURE;
    }
    const uint32_t polarity = (route == FJ_ROUTE_POSITIVE_DIRECT || route == FJ_ROUTE_POSITIVE_PRINT) ? 1u : 0u;
    if (!compose_visual_grain(in.grain_coefficients, polarity, in.grain, d.grain, diagnostic)) {
        if (diagnostic.empty()) diagnostic = kMismatch;
        return FJ_STATUS_PREPARATION_FAILURE;
    }
    if ((d.grain.flags & FJ_GRAIN_ACTIVE) != 0 &&
        (d.grain.session_seed != f.session_seed || d.grain.clip_token != f.clip_token)) {
        diagnostic = kMismatch;
        return FJ_STATUS_PREPARATION_FAILURE;
    }
    if (!effects_math_empty(in.effects)) {
        if (in.effects.present == 0) {
            diagnostic = "ResourceDescriptorMismatch phase=effects_descriptor field=rust_result";
            return FJ_STATUS_PREPARATION_FAILURE;
        }
        diagnostic = "UnsupportedInput film and gate effects are not rendered by this library";
        return FJ_STATUS_UNSUPPORTED_INPUT;
    }
    out = d;
    return FJ_STATUS_SUCCESS;
}

void auto_exposure_preview(const FjRect& source, const FjRect& meter, uint32_t method, FjAutoExposure& out) {
    constexpr int64_t kMaxLongEdge = 256;
    const int64_t w = std::max<int64_t>(0, static_cast<int64_t>(meter.x2) - meter.x1);
    const int64_t h = std::max<int64_t>(0, static_cast<int64_t>(meter.y2) - meter.y1);
    int32_t pw = static_cast<int32_t>(w), ph = static_cast<int32_t>(h);
    if (std::max(w, h) > kMaxLongEdge) {
        if (w >= h) {
            pw = static_cast<int32_t>(kMaxLongEdge);
            ph = static_cast<int32_t>(std::max<int64_t>(1, (h * kMaxLongEdge + w / 2) / w));
        } else {
            ph = static_cast<int32_t>(kMaxLongEdge);
            pw = static_cast<int32_t>(std::max<int64_t>(1, (w * kMaxLongEdge + h / 2) / h));
        }
    }
    Fnv f;
    for (int32_t v : {source.x1, source.y1, source.x2, source.y2, meter.x1, meter.y1, meter.x2, meter.y2, pw, ph})
        f.put<int32_t>(v);
    f.put<uint8_t>(static_cast<uint8_t>(method));
    out = FjAutoExposure{};
    out.source_bounds = source;
    out.meter_bounds = meter;
    out.preview_width = pw;
    out.preview_height = ph;
    out.method = method;
    out.hash = f.h == 0 ? 1 : f.h;
}

}  // namespace fjx
