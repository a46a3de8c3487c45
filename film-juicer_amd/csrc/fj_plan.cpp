// fj_plan.cpp — validation of one render call, packing of prepared tables into
// device layouts, and the frame schedule on the caller's stream.
#include "fj_plan.h"
#include "fj_autoexposure.h"
#include "fj_highlight.h"
#include "fj_diffusion_host.h"
#include "fj_glare.h"
#include "fj_grain.h"
#include "fj_grain_host.h"
#include "fj_halation_host.h"
#include "fj_host_imports.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "fj_halation.h"
#include "fj_host.h"
#include "fj_spatial.h"

namespace fjx {

// FNV-1a over the little-endian bytes of 64-bit values (the reference's
// Hash::hash_uint64_values, src/Hash.h:50-56): the glare stream's seed.
uint64_t fnv1a_u64s(const uint64_t* v, int n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < 8; ++b) {
            h ^= (v[i] >> (8 * b)) & 0xFFu;
            h *= 0x100000001b3ull;
        }
    return h;
}

// ---------------------------------------------------------------------------
// hashing: 64-bit words, multiply-xorshift mixing (content identity only)
// ---------------------------------------------------------------------------
uint64_t hash_bytes(const void* data, size_t bytes, uint64_t seed) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    uint64_t h = seed ^ (0x9E3779B97F4A7C15ull * (bytes + 1));
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h ^= w * 0xBF58476D1CE4E5B9ull;
        h = (h << 27 | h >> 37) * 0x94D049BB133111EBull;
    }
    uint64_t tail = 0;
    for (size_t j = 0; i + j < bytes; ++j) tail |= static_cast<uint64_t>(p[i + j]) << (8 * j);
    h ^= tail * 0xBF58476D1CE4E5B9ull;
    h ^= h >> 31;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 32;
    return h ? h : 1;
}

// ---------------------------------------------------------------------------
// device state
// ---------------------------------------------------------------------------
hipError_t DeviceState::init(int dev) {
    device = dev;
    hipError_t e = hipDeviceGetAttribute(&compute_units, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipMalloc(reinterpret_cast<void**>(&error_words), sizeof(unsigned int) * kRing);
    if (e != hipSuccess) return e;
    e = hipHostMalloc(reinterpret_cast<void**>(&error_host), sizeof(unsigned int) * kRing, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    for (int i = 0; i < kRing; ++i) {
        error_host[i] = 0;
        e = hipEventCreateWithFlags(&error_events[i], hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    for (int i = 0; i <= kMaxStages; ++i) {
        e = hipEventCreate(&stage_events[i]);
        if (e != hipSuccess) return e;
    }
    for (WorkspaceSlot& w : slots) {
        e = hipEventCreateWithFlags(&w.done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    for (hipEvent_t& b : batch_events) {
        e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    e = fft_kernels_init(&max_line);
    if (e != hipSuccess) return e;
    if (halation_kernels_init() != hipSuccess) {
        // Without the larger LDS window the row pass holds shorter rows.
        (void)hipGetLastError();
        halation_max_line = 7680;
    }
    return hipSuccess;
}

hipError_t DeviceState::release() {
    hipError_t first = hipSuccess;
    auto keep = [&](hipError_t e) {
        if (first == hipSuccess && e != hipSuccess) first = e;
    };
    keep(hipSetDevice(device));
    keep(hipDeviceSynchronize());
    for (auto& kv : tables) keep(hipFree(kv.second.ptr));
    tables.clear();
    transfers.clear();
    lru.clear();
    table_bytes = 0;
    for (WorkspaceSlot& w : slots) {
        if (w.ptr) keep(hipFree(w.ptr));
        if (w.done) keep(hipEventDestroy(w.done));
        w = WorkspaceSlot{};
    }
    active_slot = 0;
    for (hipStream_t& st : workers) {
        if (st) keep(hipStreamDestroy(st));
        st = nullptr;
    }
    for (hipEvent_t& b : batch_events) {
        if (b) keep(hipEventDestroy(b));
        b = nullptr;
    }
    for (int i = 0; i < kRing; ++i)
        if (error_events[i]) keep(hipEventDestroy(error_events[i]));
    for (int i = 0; i <= kMaxStages; ++i)
        if (stage_events[i]) keep(hipEventDestroy(stage_events[i]));
    for (int i = 0; i < kFitRing; ++i) {
        if (fit_events[i]) keep(hipEventDestroy(fit_events[i]));
        fit_events[i] = nullptr;
        fit_pending[i] = false;
    }
    if (fit_ring) keep(hipFree(fit_ring));
    if (fit_staging) keep(hipHostFree(fit_staging));
    fit_ring = nullptr;
    fit_staging = nullptr;
    if (error_words) keep(hipFree(error_words));
    if (error_host) keep(hipHostFree(error_host));
    error_words = nullptr;
    error_host = nullptr;
    return first;
}

hipError_t DeviceState::table(uint64_t key, const void* host, size_t bytes, const void** out) {
    auto it = tables.find(key);
    if (it != tables.end()) {
        lru.remove(key);
        lru.push_front(key);
        it->second.serial = frame_serial;
        *out = it->second.ptr;
        return hipSuccess;
    }
    constexpr size_t kBudget = size_t(2) << 30;  // device bytes kept for cached tables
    constexpr size_t kMaxEntries = 256;
    while (!lru.empty() && (table_bytes + bytes > kBudget || tables.size() >= kMaxEntries)) {
        const uint64_t victim = lru.back();
        auto v = tables.find(victim);
        // Tables the current frame already uses stay: the budget is exceeded instead.
        if (v != tables.end() && v->second.serial == frame_serial) break;
        lru.pop_back();
        if (v != tables.end()) {
            // hipFree synchronizes the device, so no in-flight frame still reads it.
            const hipError_t e = hipFree(v->second.ptr);
            table_bytes -= v->second.bytes;
            tables.erase(v);
            transfers.clear();  // they may point into the freed table
            if (e != hipSuccess) return e;
        }
    }
    DeviceTable t;
    t.bytes = bytes;
    t.serial = frame_serial;
    hipError_t e = hipMalloc(&t.ptr, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return e;
    // Synchronous copy: the prepared spans expire when the call returns. A
    // table without host bytes is filled by device work on the frame's stream.
    if (host) e = hipMemcpy(t.ptr, host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t.ptr);
        return e;
    }
    tables.emplace(key, t);
    lru.push_front(key);
    table_bytes += bytes;
    *out = t.ptr;
    return hipSuccess;
}

hipError_t DeviceState::acquire_workspace(hipStream_t stream) {
    WorkspaceSlot& w = slots[active_slot];
    if (!w.used || w.stream == stream) return hipSuccess;
    return hipStreamWaitEvent(stream, w.done, 0);
}

hipError_t DeviceState::release_workspace(hipStream_t stream) {
    WorkspaceSlot& w = slots[active_slot];
    w.stream = stream;
    w.used = true;
    return hipEventRecord(w.done, stream);
}

hipError_t DeviceState::ensure_workspace(size_t bytes) {
    WorkspaceSlot& w = slots[active_slot];
    if (w.bytes >= bytes) return hipSuccess;
    if (w.ptr) {
        // hipFree waits for the device, so no queued user of the old
        // allocation is still running.
        const hipError_t e = hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
        if (e != hipSuccess) return e;
    }
    const hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e == hipSuccess) w.bytes = bytes;
    return e;
}

unsigned int DeviceState::drain_errors(bool wait_all) {
    unsigned int bits = 0u;
    for (int i = 0; i < kRing; ++i) {
        if (!error_pending[i]) continue;
        if (wait_all) {
            if (hipEventSynchronize(error_events[i]) != hipSuccess) continue;
        } else if (hipEventQuery(error_events[i]) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        bits |= error_host[i];
        error_pending[i] = false;
    }
    return bits;
}

// ---------------------------------------------------------------------------
// validation
// ---------------------------------------------------------------------------
namespace {

struct Diag {
    std::string text;
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT;
    bool malformed(const char* field) {
        text = std::string("MalformedPreparedHostData field=") + field;
        return false;
    }
    bool unsupported(const char* what) {
        text = std::string("UnsupportedInput ") + what;
        return false;
    }
    // A descriptor the reference admits but cannot prepare.
    bool preparation(const char* what) {
        text = what;
        category = FJ_STATUS_PREPARATION_FAILURE;
        return false;
    }
};

template <typename Span>
bool exact(const Span& s, size_t count, const char* field, Diag& d) {
    return (s.count == count && ((count == 0) == (s.data == nullptr))) || d.malformed(field);
}

bool is_print_route(uint32_t route) { return route == FJ_ROUTE_NEGATIVE_PRINT || route == FJ_ROUTE_POSITIVE_PRINT; }

// native/juicer_cuda_descriptors.cpp:522-600 tag checks and
// native/juicer_cuda_prepared.cpp:58-158 span checks.
bool validate_prepared(const FjPreparedHostData& p, Diag& d) {
    if (p.route > FJ_ROUTE_POSITIVE_PRINT || p.capture_polarity > FJ_POLARITY_POSITIVE)
        return d.malformed("route_or_polarity");
    const uint32_t expected_polarity = p.route <= FJ_ROUTE_NEGATIVE_PRINT ? FJ_POLARITY_NEGATIVE : FJ_POLARITY_POSITIVE;
    if (p.capture_polarity != expected_polarity) return d.malformed("route_polarity");
    if ((p.scanner_correction.flags & ~FJ_CORRECTION_ACTIVE) != 0 || (p.scanner_post.flags & ~FJ_POST_GLARE) != 0)
        return d.malformed("scanner.flags");
    const FjOutputColor& oc = p.output_color;
    const uint32_t color_flags = FJ_COLOR_ENCODE_CCTF | FJ_COLOR_INPUT_IS_OUTPUT_SPACE | FJ_COLOR_GAMUT_COMPRESSION;
    if (oc.color_space > FJ_OUTPUT_REC709 || (oc.flags & ~color_flags) != 0) return d.malformed("output_color.tags");
    if (oc.encoding.schema != FJ_OUTPUT_ENCODING_SCHEMA || oc.encoding.cctf_kind > FJ_CCTF_DWG_INTERMEDIATE)
        return d.malformed("output_color.encoding");

    const bool print = is_print_route(p.route);
    const bool layers = (p.grain.flags & FJ_GRAIN_SUBLAYERS) != 0;
    const bool grain = (p.grain.flags & FJ_GRAIN_ACTIVE) != 0;
    const bool gamut = (oc.flags & FJ_COLOR_GAMUT_COMPRESSION) != 0;
    const FjFilmExposure& ex = p.film_exposure;
    const FjFilmDevelopment& dv = p.film_development;
    const FjDirCouplers& dir = p.dir_couplers;
    if (ex.method > FJ_RAW_ARCTIC_2026_BETA04 || ex.input_color_space > FJ_INPUT_SRGB_REC709 ||
        (ex.flags & ~(FJ_FILM_DECODE_CCTF | FJ_FILM_ADAPT_INPUT | FJ_FILM_AUTO_EXPOSURE)) != 0 ||
        dir.mode > FJ_DIR_POSITIVE_RECEIVER_LANGMUIR || p.auto_exposure.method > FJ_METER_HIGHLIGHT_WEIGHTED ||
        (p.print.flags & ~FJ_PRINT_PREFLASH) != 0)
        return d.malformed("film_or_print_tag");
    if ((!print && p.print.flags != 0) || (dir.mode == FJ_DIR_INACTIVE && (dir.hash != 0 || dir.compensated_axes_hash != 0)) ||
        (!layers && dv.density_layers_hash != 0))
        return d.malformed("inactive_family");
    if ((p.film_profile_key.count != 0) != (p.film_profile_key.data != nullptr)) return d.malformed("film_profile_key");
    if ((p.print_profile_key.count != 0) != (p.print_profile_key.data != nullptr)) return d.malformed("print_profile_key");
    const bool mallett = ex.method == FJ_RAW_MALLETT_2019;
    if (!exact(ex.sensitivity_rgb, 81 * 3, "film_exposure.sensitivity_rgb", d) ||
        !exact(ex.tc_lut_rgba, mallett ? 0 : size_t(192) * 192 * 4, "film_exposure.tc_lut_rgba", d) ||
        !exact(ex.mallett_illuminant, mallett ? 81 : 0, "film_exposure.mallett_illuminant", d) ||
        !exact(ex.mallett_basis, mallett ? 81 * 3 : 0, "film_exposure.mallett_basis", d))
        return false;
    const size_t n = dv.log_exposure.count;
    if (n == 0 || n > static_cast<size_t>(std::numeric_limits<int>::max()) / 3) return d.malformed("film_development.log_exposure.count");
    if (!exact(dv.log_exposure, n, "film_development.log_exposure", d) ||
        !exact(dv.density_rgb, n * 3, "film_development.density_rgb", d))
        return false;
    for (const auto& layer : dv.density_layers)
        for (const auto& ch : layer)
            if (!exact(ch, layers ? n : 0, "film_development.density_layers", d)) return false;
    for (const auto& axis : dir.compensated_axes_rgb)
        if (!exact(axis, dir.mode == FJ_DIR_INACTIVE ? 0 : n, "dir_couplers.compensated_axes_rgb", d)) return false;
    const FjScannerSpectra& sp = p.scanner_spectra;
    for (const FjFloatSpan* s : {&sp.dye_c, &sp.dye_m, &sp.dye_y, &sp.weighted_x, &sp.weighted_y, &sp.weighted_z})
        if (!exact(*s, 81, "scanner_spectra", d)) return false;
    if (!exact(sp.base_density, sp.base_density.count == 0 ? 0 : 81, "scanner_spectra.base_density", d) ||
        !exact(oc.gamut_cmax, gamut ? size_t(64) * 720 : 0, "output_color.gamut_cmax", d))
        return false;
    const FjPrint& pr = p.print;
    const size_t pn = print ? pr.log_exposure.count : 0;
    if (print && (pn == 0 || pn > static_cast<size_t>(std::numeric_limits<int>::max()) / 3))
        return d.malformed("print.log_exposure.count");
    const bool preflash = print && (pr.flags & FJ_PRINT_PREFLASH) != 0;
    if (!exact(pr.film_density_cmy, print ? 81 * 3 : 0, "print.film_density_cmy", d) ||
        !exact(pr.film_base_density, print ? 81 : 0, "print.film_base_density", d) ||
        !exact(pr.sensitivity_cmy, print ? 81 * 3 : 0, "print.sensitivity_cmy", d) ||
        !exact(pr.log_exposure, pn, "print.log_exposure", d) || !exact(pr.density_cmy, pn * 3, "print.density_cmy", d) ||
        !exact(pr.main_illuminant, print ? 81 : 0, "print.main_illuminant", d) ||
        !exact(pr.preflash_illuminant, preflash ? 81 : 0, "print.preflash_illuminant", d))
        return false;
    const FjStaticNoise& nz = p.noise;
    if (!grain && (nz.stbn_width || nz.stbn_height || nz.stbn_frames || nz.wang_width || nz.wang_height ||
                   nz.wang_tile_count || nz.wang_colors))
        return d.malformed("inactive_noise");
    size_t stbn = 0, wang = 0, lut = 0;
    if (grain) {
        if (nz.stbn_width <= 0 || nz.stbn_height <= 0 || nz.stbn_frames <= 0 || nz.wang_width <= 0 ||
            nz.wang_height <= 0 || nz.wang_colors <= 0 || nz.wang_tile_count == 0 || nz.wang_colors > 255 ||
            nz.wang_tile_count > static_cast<size_t>(std::numeric_limits<int>::max()))
            return d.malformed("noise.dimensions");
        stbn = size_t(nz.stbn_width) * size_t(nz.stbn_height) * size_t(nz.stbn_frames);
        wang = size_t(nz.wang_width) * size_t(nz.wang_height) * nz.wang_tile_count;
        lut = size_t(nz.wang_colors) * nz.wang_colors * nz.wang_colors * nz.wang_colors;
    }
    return exact(nz.stbn, stbn, "noise.stbn", d) && exact(nz.wang_tiles, wang, "noise.wang_tiles", d) &&
           exact(nz.wang_lut, lut, "noise.wang_lut", d);
}

uint64_t norm_key(uint64_t v) { return v ? v : 1; }

// native/juicer_cuda_prepared.cpp:160-222 (identity subset; auto exposure is
// rejected separately).
bool validate_binding(const FjPreparedHostData& p, const FjFrame& f, const FjSubmission& s, Diag& d) {
    const double floor_t = std::floor(f.time_frames);
    if (p.recipe_hash == 0 || s.instance_token == 0 || s.submission_id == 0 ||
        s.upload_core_hash != norm_key(p.recipe_hash) || s.dir_hash != norm_key(p.dir_couplers.hash) ||
        s.scanner_hash == 0 || s.auto_exposure_hash != norm_key(p.auto_exposure.hash) || !std::isfinite(f.time_frames) ||
        f.time_frames < -0x1p63 || f.time_frames >= 0x1p63 ||
        s.frame_token != static_cast<uint64_t>(static_cast<int64_t>(floor_t)))
        return d.malformed("submission_identity");
    if (p.scanner_bounds.hash == 0 || p.scanner_lut.hash == 0 || p.scanner_lut.density_bounds_hash != p.scanner_bounds.hash)
        return d.malformed("scanner_bounds.identity");
    return true;
}

bool rect_inside(const FjRect& inner, const FjRect& outer) {
    return inner.x1 >= outer.x1 && inner.y1 >= outer.y1 && inner.x2 <= outer.x2 && inner.y2 <= outer.y2;
}

bool validate_frame(const FjFrame& f, Diag& d) {
    const FjImage& s = f.source;
    const FjImage& o = f.destination;
    if (s.depth != FJ_DEPTH_FLOAT32 || o.depth != FJ_DEPTH_FLOAT32) return d.unsupported("image depth (float32 only)");
    if ((s.components != FJ_COMPONENTS_RGB && s.components != FJ_COMPONENTS_RGBA) || o.components != s.components)
        return d.unsupported("image components (RGB or RGBA, equal on both images)");
    const FjRect& w = f.render_window;
    if (w.x2 <= w.x1 || w.y2 <= w.y1) return d.unsupported("empty render window");
    if (!rect_inside(w, s.bounds) || !rect_inside(w, o.bounds)) return d.unsupported("render window outside image bounds");
    const long long pixel_bytes = 4LL * s.components;
    if (s.row_bytes < pixel_bytes * (s.bounds.x2 - s.bounds.x1) || o.row_bytes < pixel_bytes * (o.bounds.x2 - o.bounds.x1) ||
        s.row_bytes % 4 != 0 || o.row_bytes % 4 != 0)
        return d.unsupported("image row_bytes");
    if (s.components == 4 && ((s.address | o.address | static_cast<uintptr_t>(s.row_bytes) |
                               static_cast<uintptr_t>(o.row_bytes)) & 15u) != 0)
        return d.unsupported("RGBA images must be 16-byte aligned");
    return true;
}

// Scanner Gaussian kernels: a bound post descriptor prepares its lens-blur
// and unsharp kernels whenever their sigma is positive, and refuses a
// non-finite sigma or a radius outside 1..75 (prepare_scanner_post_effects,
// src/ProcessRoot.cpp:3121-3158; ensure_gaussian_kernel,
// src/Cuda/JuicerCudaResources.cpp:6436-6448).
constexpr int kMaxGaussianRadius = 75;

bool gaussian_valid(float sigma, int radius) {
    return std::isfinite(sigma) && sigma > 0.0f && radius >= 1 && radius <= kMaxGaussianRadius;
}

// Scanner glare is generated when the post descriptor is bound and its
// percent is finite and positive with a finite roughness
// (JuicerCudaScanPipeline.cu:2361-2364, JuicerCudaExecutor.cpp:2405-2413),
// and blurred with the Gaussian that the glare flag prepares
// (ProcessRoot.cpp:3147-3154).
bool glare_active(const FjScannerPostEffects& sp) {
    return sp.hash != 0 && std::isfinite(sp.glare_percent) && sp.glare_percent > 0.0f &&
           std::isfinite(sp.glare_roughness);
}

bool glare_blurred(const FjScannerPostEffects& sp) {
    return (sp.flags & FJ_POST_GLARE) != 0 && sp.glare_blur_sigma_px > 0.0f;
}

bool post_kernels_valid(const FjScannerPostEffects& sp) {
    if (sp.hash == 0) return true;
    if (sp.lens_blur_sigma_px > 0.0f && !gaussian_valid(sp.lens_blur_sigma_px, sp.lens_blur_radius)) return false;
    if (sp.unsharp_sigma_px > 0.0f && !gaussian_valid(sp.unsharp_sigma_px, sp.unsharp_radius)) return false;
    if (glare_blurred(sp) && !gaussian_valid(sp.glare_blur_sigma_px, sp.glare_blur_radius)) return false;
    return true;
}

// Launch conditions (launch_focused_scanner_post_output,
// JuicerCudaScanPipeline.cu:2793-2813): lens blur with a prepared kernel;
// the unsharp mask with a prepared kernel and a finite positive amount.
bool lens_active(const FjScannerPostEffects& sp) { return sp.hash != 0 && sp.lens_blur_sigma_px > 0.0f; }

bool unsharp_active(const FjScannerPostEffects& sp) {
    return sp.hash != 0 && sp.unsharp_sigma_px > 0.0f && sp.unsharp_radius > 0 && std::isfinite(sp.unsharp_amount) &&
           sp.unsharp_amount > 0.0f;
}

// Post effects on the scanner-linear planes: lens blur, then the unsharp
// mask.
bool post_active(const FjScannerPostEffects& sp) { return unsharp_active(sp) || lens_active(sp); }

// Unsharp masks wider than the fused post kernel's LDS tile run as a
// separate blur + combine before the encode.
bool unsharp_separate(const FjScannerPostEffects& sp) { return unsharp_active(sp) && sp.unsharp_radius > max_post_radius(); }

}  // namespace

bool highlight_boost_valid(const FjHighlightBoost& hb) {
    // The recipe's sharpness is 28^(1 - range), range in [0, 1]
    // (rust optics/highlight_boost.rs:56, :84); larger shapes overflow the
    // curve's normalization.
    return std::isfinite(hb.protection) && std::isfinite(hb.sharpness) && std::isfinite(hb.peak_gain) &&
           hb.protection >= 0.0 && hb.sharpness >= 0.0 && hb.sharpness <= kHighlightMaxSharpness && hb.peak_gain >= 1.0;
}

namespace {

// Families this build renders; everything else is rejected loudly.
bool check_supported(const FjPreparedHostData& p, const FjFrame& f, Diag& d) {
    if (p.film_exposure.flags & FJ_FILM_AUTO_EXPOSURE) {
        // meter_camera_exposure (JuicerCudaExecutor.cpp:1013-1066) and the
        // launcher's checks (JuicerCudaAutoExposure.cu:700-740, :3145-3147).
        const FjAutoExposure& a = p.auto_exposure;
        const FjRect& sb = f.source.bounds;
        if (a.preview_width <= 0 || a.preview_height <= 0)
            return d.unsupported("MissingRequiredResource phase=4C field=auto_exposure_buffers");
        if (a.preview_width > 256 || a.preview_height > 256 || a.meter_bounds.x2 <= a.meter_bounds.x1 ||
            a.meter_bounds.y2 <= a.meter_bounds.y1)
            return d.malformed("auto_exposure.preview");
        if (a.source_bounds.x1 != sb.x1 || a.source_bounds.y1 != sb.y1 || a.source_bounds.x2 != sb.x2 ||
            a.source_bounds.y2 != sb.y2)
            return d.unsupported("auto exposure source bounds differ from the source image");
        const bool print_route = p.route == FJ_ROUTE_NEGATIVE_PRINT || p.route == FJ_ROUTE_POSITIVE_PRINT;
        if (!print_route && a.method == FJ_METER_MEDIAN) return d.unsupported("QuantizedMedianNotAcceptedForPhase3");
    }
    if (!post_kernels_valid(p.scanner_post)) return d.preparation("gaussian descriptor is invalid");
    if (p.optics.flags & ~(FJ_OPTICS_CAMERA_DIFFUSION | FJ_OPTICS_ENLARGER_DIFFUSION | FJ_OPTICS_SCATTER_HALATION |
                           FJ_OPTICS_HIGHLIGHT_BOOST))
        return d.malformed("optics.flags");
    {
        std::string field;
        if (!diffusion_decode(p.optics, is_print_route(p.route), field)) return d.malformed(field.c_str());
    }
    {
        // Highlight boost runs when its recipe is bound
        // (launch_highlight_boost, JuicerCudaHighlightBoost.cu:195-205).
        const FjHighlightBoost& hb = p.optics.highlight_boost;
        const bool flag = (p.optics.flags & FJ_OPTICS_HIGHLIGHT_BOOST) != 0;
        if (flag != (hb.recipe_hash != 0)) return d.malformed("optics.highlight_boost");
        if (flag && !highlight_boost_valid(hb)) return d.malformed("optics.highlight_boost");
    }
    if (p.optics.flags & FJ_OPTICS_SCATTER_HALATION) {
        std::string field;
        if (!halation_validate(p.optics.scatter_halation, field)) return d.malformed(field.c_str());
    } else if (p.optics.scatter_halation.recipe_hash != 0) {
        return d.malformed("optics.inactive_scatter_halation");
    }
    if (p.optics.flags != 0) {
        // Optics need the complete frame domain (native/juicer_cuda_api.cpp:432-436).
        auto same = [](const FjRect& a, const FjRect& b) {
            return a.x1 == b.x1 && a.y1 == b.y1 && a.x2 == b.x2 && a.y2 == b.y2;
        };
        if (!same(f.source.bounds, f.full_frame_extent) || !same(f.destination.bounds, f.full_frame_extent) ||
            !same(f.render_window, f.full_frame_extent))
            return d.unsupported("optics require the complete frame domain");
    }
    {
        std::string field;
        if (!grain_decode(p.grain, field)) return d.malformed(field.c_str());
    }
    {
        // Effects descriptor (native/juicer_cuda_descriptors.cpp:466-482);
        // of its families this build renders the gate weave alone.
        const FjFilmEffects& fx = p.effects;
        if (fx.flags & ~(FJ_EFFECTS_WEAVE | FJ_EFFECTS_FILM | FJ_EFFECTS_GATE_TRANSMITTANCE | FJ_EFFECTS_GATE_OUTPUT |
                         FJ_EFFECTS_FULL_FRAME))
            return d.malformed("effects.flags");
        if ((fx.flags & (FJ_EFFECTS_FILM | FJ_EFFECTS_GATE_OUTPUT)) == 0)
            return (fx.flags == 0 && fx.hash == 0 && fx.recipe_hash == 0) || d.malformed("effects.inactive");
        if (fx.hash == 0 || fx.recipe_hash == 0) return d.malformed("effects.identity");
        if (fx.flags & (FJ_EFFECTS_FILM | FJ_EFFECTS_GATE_TRANSMITTANCE))
            return d.unsupported("film/gate dust and scratches");
        if (fx.flags & FJ_EFFECTS_FULL_FRAME) {
            // film_juicer_effects_full_frame_preflight (JuicerCudaExecutor.cpp:751-769).
            const FjExtent& e = fx.full_frame_extent;
            auto full = [&e](const FjRect& r) {
                return r.x1 == e.x && r.y1 == e.y && r.x2 - r.x1 == e.width && r.y2 - r.y1 == e.height;
            };
            if (!full(f.render_window) || !full(f.source.bounds) || !full(f.full_frame_extent))
                return d.unsupported("film effects require the complete frame domain");
        }
    }
    return true;
}

// The gate weave runs with gate output and the weave flag, unless a grain
// debug view replaces the image (JuicerCudaExecutor.cpp:2582-2589, :1769-1773).
bool weave_active(const FjPreparedHostData& p) {
    const bool gate_output = (p.effects.flags & FJ_EFFECTS_GATE_OUTPUT) != 0 && p.effects.hash != 0;
    const bool grain_debug = (p.grain.flags & FJ_GRAIN_ACTIVE) != 0 && p.grain.debug_view != 0;
    return gate_output && (p.effects.flags & FJ_EFFECTS_WEAVE) != 0 && !grain_debug;
}

bool same_extent(const FjExtent& e, const FjRect& r) {
    return e.x == r.x1 && e.y == r.y1 && e.width == r.x2 - r.x1 && e.height == r.y2 - r.y1;
}

// Spatial DIR binding (native/juicer_cuda_prepared.cpp:192-199).
bool check_spatial_dir(const FjPreparedHostData& p, const FjFrame& f, int max_line, Diag& d) {
    const FjSpatialDir& sd = p.spatial_dir;
    if (sd.hash == 0) return sd.component_count == 0 || d.malformed("spatial_dir.hash");
    if (p.dir_couplers.mode == FJ_DIR_INACTIVE) return d.malformed("spatial_dir.inactive_dir");
    if (sd.recipe_hash != p.dir_couplers.hash || !same_extent(sd.render_extent, f.render_window) ||
        !same_extent(sd.render_extent, f.full_frame_extent) || !same_extent(sd.filter_domain_extent, f.full_frame_extent) ||
        !same_extent(sd.full_frame_extent, f.full_frame_extent))
        return d.malformed("spatial_dir.frame");
    const int w = f.render_window.x2 - f.render_window.x1, h = f.render_window.y2 - f.render_window.y1;
    // Each side's line transform (a Bluestein plan's m-point one for a side
    // that is not 2,3,5-smooth) must fit the LDS engine.
    FftPlanDev plan_w, plan_h;
    if (!make_fft_plan(w, plan_w) || !make_fft_plan(h, plan_h) || fft_length(plan_w) > max_line ||
        fft_length(plan_h) > max_line)
        return d.unsupported("spatial DIR frame side above the LDS line limit");
    // A host that defines the boundary functions certifies each recursive
    // component's reflected boundary on this frame.
    uint32_t category = 0;
    std::string diag;
    if (!host_dir_boundaries(sd, w, h, category, diag)) {
        d.category = category;
        d.text = diag;
        return false;
    }
    return true;
}

// Diffusion binding: the optics frame is the whole frame
// (native/juicer_cuda_prepared.cpp:216-221).
bool check_optics_frame(const FjPreparedHostData& p, const FjFrame& f, Diag& d) {
    if (p.optics.diffusion_hash == 0 && p.optics.highlight_boost.recipe_hash == 0) return true;
    if (!same_extent(p.optics.full_frame, f.source.bounds) || !same_extent(p.optics.full_frame, f.render_window) ||
        !same_extent(p.optics.full_frame, f.full_frame_extent))
        return d.malformed("optics.frame");
    return true;
}

// Grain binding (native/juicer_cuda_prepared.cpp:200-207); a descriptor whose
// blurs need the whole frame renders the whole frame.
bool check_grain_frame(const FjPreparedHostData& p, const FjFrame& f, Diag& d) {
    if (!grain_binding(p.grain, f)) return d.malformed("grain.frame");
    if ((p.grain.flags & FJ_GRAIN_FULL_FRAME) != 0) {
        const FjRect& w = f.render_window;
        const FjRect& full = f.full_frame_extent;
        if (w.x1 != full.x1 || w.y1 != full.y1 || w.x2 != full.x2 || w.y2 != full.y2)
            return d.unsupported("grain blurs require the complete frame domain");
    }
    return true;
}


// ---------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------
void mat_mul(const float* a, const float* b, float* out) {  // out = a * b (double accumulation)
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += static_cast<double>(a[r * 3 + k]) * b[k * 3 + c];
            out[r * 3 + c] = static_cast<float>(s);
        }
}

// Film curves [x | yB | yG | yR | cxB | cxG | cxR] (JuicerCudaResources.cpp:3724-3842).
std::vector<float> pack_film_curves(const FjPreparedHostData& p) {
    const FjFilmDevelopment& dv = p.film_development;
    const size_t n = dv.log_exposure.count;
    const bool dir = p.dir_couplers.mode != FJ_DIR_INACTIVE;
    std::vector<float> out((dir ? 7 : 4) * n);
    std::memcpy(out.data(), dv.log_exposure.data, n * sizeof(float));
    for (size_t i = 0; i < n; ++i)
        for (int b = 0; b < 3; ++b) out[(1 + b) * n + i] = dv.density_rgb.data[i * 3 + (2 - b)];
    if (dir)
        for (int b = 0; b < 3; ++b) std::memcpy(&out[(4 + b) * n], p.dir_couplers.compensated_axes_rgb[2 - b].data, n * sizeof(float));
    return out;
}

// Compacted print spectra (JuicerCudaDeviceHelpers.cuh:866-933): wavelengths
// whose transmitted light is NaN for every density (NaN extinction, baseline
// or illuminant) or whose three sensitivities are NaN contribute nothing and
// are dropped; the -log2(10) factor and the illuminant are folded in.
std::vector<float> pack_print_spectra(const FjPrint& pr, int& k_out) {
    std::vector<float> out;
    out.reserve(81 * 8);
    const double kNegLog2_10 = -3.32192809488736234787;
    int k = 0;
    for (int i = 0; i < 81; ++i) {
        const float* e = pr.film_density_cmy.data + i * 3;
        const float* s = pr.sensitivity_cmy.data + i * 3;
        const float base = pr.film_base_density.data[i];
        const float ill = pr.main_illuminant.data[i];
        const bool any_sens = !std::isnan(s[0]) || !std::isnan(s[1]) || !std::isnan(s[2]);
        if (!any_sens || std::isnan(e[0]) || std::isnan(e[1]) || std::isnan(e[2]) || std::isnan(base) || std::isnan(ill))
            continue;
        out.push_back(static_cast<float>(e[0] * kNegLog2_10));
        out.push_back(static_cast<float>(e[1] * kNegLog2_10));
        out.push_back(static_cast<float>(e[2] * kNegLog2_10));
        out.push_back(static_cast<float>(base * kNegLog2_10));
        for (int c = 0; c < 3; ++c) out.push_back(std::isnan(s[c]) ? 0.0f : static_cast<float>(double(s[c]) * ill));
        out.push_back(0.0f);
        ++k;
    }
    k_out = k;
    if (out.empty()) out.resize(8, 0.0f);
    return out;
}

// Mallett table (JuicerCudaDeviceHelpers.cuh:619-668): wavelengths with a
// non-finite basis or illuminant never contribute and are dropped; the
// illuminant is folded into the basis, NaN sensitivities become zero.
std::vector<float> pack_mallett(const FjFilmExposure& ex, int& k_out) {
    std::vector<float> out;
    int k = 0;
    for (int i = 0; i < 81; ++i) {
        const float* b = ex.mallett_basis.data + i * 3;
        const float ill = ex.mallett_illuminant.data[i];
        const float* s = ex.sensitivity_rgb.data + i * 3;
        if (!std::isfinite(b[0]) || !std::isfinite(b[1]) || !std::isfinite(b[2]) || !std::isfinite(ill)) continue;
        for (int c = 0; c < 3; ++c) out.push_back(static_cast<float>(double(b[c]) * ill));
        out.push_back(0.0f);
        for (int c = 2; c >= 0; --c) out.push_back(std::isfinite(s[c]) ? s[c] : 0.0f);
        out.push_back(0.0f);
        ++k;
    }
    k_out = k;
    if (out.empty()) out.resize(8, 0.0f);
    return out;
}

std::vector<float> pack_print_curves(const FjPrint& pr) {
    const size_t n = pr.log_exposure.count;
    std::vector<float> out(4 * n);
    std::memcpy(out.data(), pr.log_exposure.data, n * sizeof(float));
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[(1 + c) * n + i] = pr.density_cmy.data[i * 3 + c];
    return out;
}

struct PackedLut {
    std::vector<float> nodes;  // 12 floats per node
    std::vector<float> cells;  // 8 floats per cell
    int res = 0;
};

bool pack_scanner_lut(const FjPreparedHostData& p, PackedLut& out, std::string& diag) {
    const bool print = is_print_route(p.route);
    ScannerLutInput in{};
    const FjScannerSpectra& s = p.scanner_spectra;
    in.eps_c = s.dye_c.data;
    in.eps_m = s.dye_m.data;
    in.eps_y = s.dye_y.data;
    in.weighted_x = s.weighted_x.data;
    in.weighted_y = s.weighted_y.data;
    in.weighted_z = s.weighted_z.data;
    in.baseline = s.base_density.count ? s.base_density.data : nullptr;
    in.sample_count = 81;
    in.inv_yn = s.inverse_y_normalization;
    // Scanner medium and range as packed by JuicerCudaResources.cpp:3914-3931.
    for (int c = 0; c < 3; ++c) {
        in.range_min_cmy[c] = print ? p.scanner_bounds.min_cmy[c] : -p.scanner_bounds.min_cmy[c];
        in.inverse_span_cmy[c] = p.scanner_bounds.inv_span_cmy[c];
    }
    in.medium = print ? 1 : 0;
    in.resolution = static_cast<int>(p.scanner_lut.resolution);
    ScannerLut lut;
    if (!frame_scanner_lut(in, lut, diag)) return false;
    const int r = lut.resolution;
    const size_t nodes = size_t(r) * r * r;
    out.res = r;
    out.nodes.resize(nodes * 12);
    for (size_t i = 0; i < nodes; ++i) {
        float* d = &out.nodes[i * 12];
        for (int o = 0; o < 3; ++o) {
            d[o] = lut.log2_xyz[i * 3 + o];
            d[3 + o] = lut.slope_c[i * 3 + o];
            d[6 + o] = lut.slope_m[i * 3 + o];
            d[9 + o] = lut.slope_y[i * 3 + o];
        }
    }
    const size_t cells = size_t(r - 1) * (r - 1) * (r - 1);
    out.cells.assign(cells * 8, 0.0f);
    for (size_t i = 0; i < cells; ++i) {
        float* d = &out.cells[i * 8];
        for (int o = 0; o < 3; ++o) {
            d[o] = lut.cell_min[i * 3 + o];
            d[3 + o] = lut.cell_max[i * 3 + o];
        }
    }
    return true;
}

uint64_t span_hash(const FjFloatSpan& s, uint64_t seed) {
    return s.count ? hash_bytes(s.data, s.count * sizeof(float), seed) : hash_bytes(nullptr, 0, seed);
}

// ---------------------------------------------------------------------------
// frame parameter assembly
// ---------------------------------------------------------------------------
enum TableKind : uint64_t { kTcLut = 1, kFilmCurves, kPrintSpec, kPrintCurves, kScanNodes, kScanCells, kGamut, kMallett };

uint64_t table_key(TableKind kind, uint64_t content) { return hash_bytes(&content, sizeof(content), 0xF11F0000ull + kind); }

hipError_t assemble(DeviceState& dev, const FjPreparedHostData& p, PixelParams& pp, Diag& d, bool& ok,
                    const GivenTables* given = nullptr) {
    ok = false;
    const FjFilmExposure& ex = p.film_exposure;
    // Exposure: one input matrix (adaptation folded in) and one scale.
    if (ex.flags & FJ_FILM_ADAPT_INPUT) {
        mat_mul(ex.input_xyz_adapt, ex.input_rgb_to_xyz, pp.ex.rgb_to_xyz);
    } else {
        std::memcpy(pp.ex.rgb_to_xyz, ex.input_rgb_to_xyz, sizeof(pp.ex.rgb_to_xyz));
    }
    pp.ex.decode_cctf = (ex.flags & FJ_FILM_DECODE_CCTF) && (ex.input_color_space == FJ_INPUT_BT2020 ||
                                                          ex.input_color_space == FJ_INPUT_SRGB_REC709)
                            ? static_cast<int>(ex.input_color_space)
                            : 0;
    // JuicerCudaFilmPayloads.cpp:102-112 validates the manual scale.
    const float manual = std::exp2(ex.manual_exposure_ev);
    if (!std::isfinite(ex.manual_exposure_ev) || !std::isfinite(manual) || !(manual > 0.0f) ||
        !std::isfinite(ex.route_correction_scale) || !(ex.route_correction_scale > 0.0f)) {
        d.malformed("film_exposure_scale");
        return hipSuccess;
    }
    float scale = manual * ex.route_correction_scale;
    pp.ex.scale = std::isfinite(scale) && scale > 0.0f ? scale : 1.0f;
    pp.ex.camera_scale = manual;
    pp.ex.route_scale = ex.route_correction_scale;
    const void* ptr = nullptr;
    hipError_t e = hipSuccess;
    if (ex.method == FJ_RAW_MALLETT_2019) {
        // JuicerCudaFilmPayloads.cpp:113-121 validates the mid-gray scale.
        if (!std::isfinite(ex.mallett_green_midgray_scale) || !(ex.mallett_green_midgray_scale > 0.0f)) {
            d.malformed("mallett_midgray_scale");
            return hipSuccess;
        }
        float srgb[9];
        mat_mul(ex.xyz_to_linear_srgb, pp.ex.rgb_to_xyz, srgb);
        std::memcpy(pp.ex.rgb_to_srgb, srgb, sizeof(srgb));
        int k = 0;
        std::vector<float> tab = pack_mallett(ex, k);
        e = dev.table(table_key(kMallett, hash_bytes(tab.data(), tab.size() * 4, 13)), tab.data(), tab.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        pp.ex.mallett = static_cast<const float4*>(ptr);
        pp.ex.mallett_k = k;
        pp.ex.mallett_scale = ex.mallett_green_midgray_scale;
    } else {
        const uint64_t tc_id = ex.tc_lut_hash ? ex.tc_lut_hash : span_hash(ex.tc_lut_rgba, 11);
        e = dev.table(table_key(kTcLut, tc_id), ex.tc_lut_rgba.data, ex.tc_lut_rgba.count * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
        pp.ex.tc_lut = static_cast<const float4*>(ptr);
    }

    // Development.
    const FjFilmDevelopment& dv = p.film_development;
    if (given && given->film_curves) {
        pp.dev.curves = given->film_curves;
    } else {
        std::vector<float> film = pack_film_curves(p);
        e = dev.table(table_key(kFilmCurves, hash_bytes(film.data(), film.size() * 4, 21)), film.data(),
                      film.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        pp.dev.curves = static_cast<const float*>(ptr);
    }
    pp.dev.n = static_cast<int>(dv.log_exposure.count);
    for (int b = 0; b < 3; ++b) pp.dev.gamma_bgr[b] = dv.gamma_rgb[2 - b];
    const FjDirCouplers& dir = p.dir_couplers;
    pp.dev.dir_mode = static_cast<int>(dir.mode);
    for (int db = 0; db < 3; ++db) {
        for (int rb = 0; rb < 3; ++rb) pp.dev.M[db * 3 + rb] = dir.matrix_rgb[(2 - db) * 3 + (2 - rb)];
        pp.dev.dmax[db] = dir.density_max_rgb[2 - db];
        pp.dev.dref[db] = dir.density_ref_rgb[2 - db];
        pp.dev.donor_k[db] = dir.donor_k_rgb[2 - db];
        pp.dev.recv_cref[db] = dir.receiver_c_ref_rgb[2 - db];
        pp.dev.recv_kr[db] = dir.receiver_kr_rgb[2 - db];
    }

    // Print (JuicerCudaResources.cpp:4331-4420 scale checks).
    pp.print = is_print_route(p.route) ? 1 : 0;
    if (pp.print) {
        const FjPrint& pr = p.print;
        const float route_scale = p.scanner_correction.exposure_scale;
        if (!std::isfinite(pr.exposure) || pr.exposure < 0.0f || !std::isfinite(pr.normalizer) || !(pr.normalizer > 0.0f) ||
            !std::isfinite(route_scale) || !(route_scale > 0.0f)) {
            d.malformed("print_exposure_scale");
            return hipSuccess;
        }
        const bool preflash = (pr.flags & FJ_PRINT_PREFLASH) != 0;
        if (preflash) {
            if (!std::isfinite(pr.preflash_exposure) || pr.preflash_exposure < 0.0f) {
                d.malformed("preflash_exposure");
                return hipSuccess;
            }
            for (float raw : pr.preflash_raw_cmy)
                if (!std::isfinite(raw) || raw < 0.0f) {
                    d.malformed("preflash_raw");
                    return hipSuccess;
                }
        }
        if (given && given->print_spectra) {
            pp.pr.spec = given->print_spectra;
            pp.pr.k = given->print_k;
        } else {
            int k = 0;
            std::vector<float> spec = pack_print_spectra(pr, k);
            e = dev.table(table_key(kPrintSpec, hash_bytes(spec.data(), spec.size() * 4, 31)), spec.data(),
                          spec.size() * 4, &ptr);
            if (e != hipSuccess) return e;
            pp.pr.spec = static_cast<const float4*>(ptr);
            pp.pr.k = k;
        }
        pp.pr.normalizer = pr.normalizer;
        const float pf = preflash && pr.preflash_exposure > 0.0f ? pr.preflash_exposure : 0.0f;
        for (int c = 0; c < 3; ++c) pp.pr.preflash[c] = pf > 0.0f ? pr.preflash_raw_cmy[c] * pf : 0.0f;
        pp.pr.outer_scale = pr.exposure * route_scale;
        std::vector<float> pc = pack_print_curves(pr);
        e = dev.table(table_key(kPrintCurves, hash_bytes(pc.data(), pc.size() * 4, 41)), pc.data(), pc.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        pp.pr.curves = static_cast<const float*>(ptr);
        pp.pr.n = static_cast<int>(pr.log_exposure.count);
    }

    // Scanner LUT: built natively, cached by the prepared scanner identity.
    uint64_t lut_id = p.scanner_lut.hash;
    {
        const FjScannerSpectra& s = p.scanner_spectra;
        uint64_t h = lut_id;
        for (const FjFloatSpan* sp : {&s.dye_c, &s.dye_m, &s.dye_y, &s.weighted_x, &s.weighted_y, &s.weighted_z, &s.base_density})
            h = span_hash(*sp, h);
        h = hash_bytes(&p.scanner_bounds, sizeof(float) * 6, h);
        const uint32_t extra[3] = {p.scanner_lut.resolution, static_cast<uint32_t>(is_print_route(p.route)),
                                   *reinterpret_cast<const uint32_t*>(&s.inverse_y_normalization)};
        lut_id = hash_bytes(extra, sizeof(extra), h);
    }
    const uint64_t node_key = table_key(kScanNodes, lut_id), cell_key = table_key(kScanCells, lut_id);
    if (dev.tables.count(node_key) == 0 || dev.tables.count(cell_key) == 0) {
        PackedLut lut;
        std::string diag;
        if (!pack_scanner_lut(p, lut, diag)) {
            uint32_t category = 0;
            std::string text;
            if (take_host_failure(category, text)) {
                d.category = category;
                d.text = text;
            } else {
                d.category = FJ_STATUS_PREPARATION_FAILURE;
                d.text = "PreparationFailure " + diag;
            }
            return hipSuccess;
        }
        e = dev.table(node_key, lut.nodes.data(), lut.nodes.size() * 4, &ptr);
        if (e != hipSuccess) return e;
        e = dev.table(cell_key, lut.cells.data(), lut.cells.size() * 4, &ptr);
        if (e != hipSuccess) return e;
    }
    e = dev.table(node_key, nullptr, 0, &ptr);
    if (e != hipSuccess) return e;
    pp.sc.nodes = static_cast<const float4*>(ptr);
    e = dev.table(cell_key, nullptr, 0, &ptr);
    if (e != hipSuccess) return e;
    pp.sc.cells = static_cast<const float4*>(ptr);
    pp.sc.res = static_cast<int>(p.scanner_lut.resolution);
    for (int c = 0; c < 3; ++c) {
        pp.sc.dmin[c] = p.scanner_bounds.min_cmy[c];
        pp.sc.inv_span[c] = p.scanner_bounds.inv_span_cmy[c];
    }
    pp.sc.correction = (p.scanner_correction.flags & FJ_CORRECTION_ACTIVE) != 0;
    pp.sc.corr_slope = p.scanner_correction.xyz_slope;
    pp.sc.corr_offset = p.scanner_correction.xyz_offset;
    const FjOutputColor& oc = p.output_color;
    mat_mul(oc.xyz_to_rgb, oc.cat02, pp.sc.xyz_to_rgb);

    // Colour.
    pp.co.gamut = (oc.flags & FJ_COLOR_GAMUT_COMPRESSION) != 0;
    if (pp.co.gamut) {
        const uint64_t g_id = oc.gamut_table_hash ? oc.gamut_table_hash : span_hash(oc.gamut_cmax, 51);
        e = dev.table(table_key(kGamut, g_id), oc.gamut_cmax.data, oc.gamut_cmax.count * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
        pp.co.cmax = static_cast<const float*>(ptr);
    }
    std::memcpy(pp.co.native_to_xyz, oc.native_rgb_to_d65_xyz, sizeof(pp.co.native_to_xyz));
    std::memcpy(pp.co.xyz_to_native, oc.d65_xyz_to_native_rgb, sizeof(pp.co.xyz_to_native));
    std::memcpy(pp.co.lknee, oc.lightness_knee, sizeof(pp.co.lknee));
    std::memcpy(pp.co.cknee, oc.chroma_knee, sizeof(pp.co.cknee));
    pp.co.encode = (oc.flags & FJ_COLOR_ENCODE_CCTF) != 0;
    pp.co.matrix = (oc.flags & FJ_COLOR_INPUT_IS_OUTPUT_SPACE) == 0;
    std::memcpy(pp.co.to_output, oc.encoding.dwg_to_output, sizeof(pp.co.to_output));
    pp.co.cctf = static_cast<int>(oc.encoding.cctf_kind);
    pp.co.gamma = oc.encoding.gamma;
    pp.co.a = oc.encoding.a;
    pp.co.b = oc.encoding.b;
    pp.co.c = oc.encoding.c;
    pp.co.d = oc.encoding.d;
    pp.co.cutoff = oc.encoding.linear_cutoff;
    ok = true;
    return hipSuccess;
}

FjStatus status_of(uint32_t category, uint32_t api = FJ_API_NONE, int32_t code = 0) {
    FjStatus s{category, api, api == FJ_API_NONE ? 0 : code};
    return s;
}

void put_error(FjErrorBuffer* error, const std::string& text) {
    if (!error || error->capacity == 0 || !error->data) {
        if (error) error->length = 0;
        return;
    }
    const size_t n = std::min(text.size(), error->capacity - 1);
    std::memcpy(error->data, text.data(), n);
    error->data[n] = '\0';
    error->length = n;
}

FjRenderOutcome outcome_hip(hipError_t e, const char* where, FjErrorBuffer* error) {
    (void)hipGetLastError();  // do not leave a sticky error for the caller's next HIP call
    FjRenderOutcome o{};
    o.status = status_of(e == hipErrorOutOfMemory ? FJ_STATUS_ALLOCATION_FAILURE : FJ_STATUS_CUDA_FAILURE,
                         FJ_API_CUDA_RUNTIME, static_cast<int32_t>(e));
    put_error(error, std::string("HIP failure in ") + where + ": " + hipGetErrorString(e));
    return o;
}

FjRenderOutcome outcome_diag(uint32_t category, const std::string& text, FjErrorBuffer* error) {
    FjRenderOutcome o{};
    o.status = status_of(category);
    put_error(error, text);
    return o;
}

// The host preparation's failure when a builder that used it failed, else
// the given one.
FjRenderOutcome host_or(uint32_t category, const std::string& text, FjErrorBuffer* error) {
    uint32_t host_category = 0;
    std::string host_text;
    if (take_host_failure(host_category, host_text)) return outcome_diag(host_category, host_text, error);
    return outcome_diag(category, text, error);
}

// check_abort (JuicerCudaExecutor.cpp:2544-2559), polled before the frame's
// preparation, after it, and after its launches are queued (:2816, :2871,
// :3125). True when the frame ends here with `out`.
bool aborted(const FjAbortCallback& cb, FjRenderOutcome& out, FjErrorBuffer* error) {
    if (!cb.query) return false;
    switch (cb.query(cb.user)) {
        case FJ_ABORT_CONTINUE:
            return false;
        case FJ_ABORT_REQUESTED:
            out = outcome_diag(FJ_STATUS_CANCELLED, "CUDA render cancelled", error);
            return true;
        default:
            out = outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "invalid CUDA abort query result", error);
            return true;
    }
}

}  // namespace

// Generates the glare field of a width x height window into `field` and,
// with `weights`, blurs it through `tmp` (optics_glare_generate_kernel and
// blur_plane_in_place, JuicerCudaScanPipeline.cu:273-314, :2361-2405).
hipError_t dir_transfer_cached(DeviceState& dev, const FjSpatialDir& sd, int width, int height, DctTransfer& out,
                               std::string& diag, bool& ok) {
    ok = false;
    const uint64_t tkey = hash_bytes(&sd.components, sizeof(sd.components),
                                     static_cast<uint64_t>(width) << 32 | static_cast<uint32_t>(height)) ^
                          (static_cast<uint64_t>(sd.component_count) << 56);
    auto cached = dev.transfers.find(tkey);
    if (cached == dev.transfers.end()) {
        HostTransfer ht;
        if (!dir_transfer(sd, width, height, ht, diag)) return hipSuccess;
        DctTransfer dt;
        const hipError_t e = upload_transfer(dev, ht, tkey, dt);
        if (e != hipSuccess) return e;
        if (dev.transfers.size() > 64) dev.transfers.clear();
        cached = dev.transfers.emplace(tkey, dt).first;
    }
    out = cached->second;
    ok = true;
    return hipSuccess;
}

bool scanner_post_active(const FjScannerPostEffects& sp) { return post_active(sp) || glare_active(sp); }

bool record_valid(const FjPreparedHostData& p, uint32_t& category, std::string& text) {
    Diag d;
    if (validate_prepared(p, d)) return true;
    category = d.category;
    text = d.text;
    return false;
}

std::vector<float> packed_film_curves(const FjPreparedHostData& p) { return pack_film_curves(p); }

std::vector<float> packed_print_spectra(const FjPrint& pr, int& k) { return pack_print_spectra(pr, k); }

hipError_t assemble_pixel_params(DeviceState& dev, const FjPreparedHostData& p, PixelParams& pp, uint32_t& category,
                                 std::string& text, bool& ok, const GivenTables* given) {
    Diag d;
    const hipError_t e = assemble(dev, p, pp, d, ok, given);
    if (!ok) {
        category = d.category;
        text = d.text;
    }
    return e;
}

static hipError_t launch_glare(DeviceState& dev, const FjScannerPostEffects& sp, const std::vector<float>& weights,
                        uint64_t seed, uint32_t medium, int origin_x, int origin_y, int width, int height, float* field,
                        float* tmp, float* uniforms, hipStream_t stream) {
    GlareArgs ga{};
    ga.out = field;
    ga.width = width;
    ga.height = height;
    ga.seed = seed;
    ga.medium = medium;
    ga.origin_x = origin_x;
    ga.origin_y = origin_y;
    ga.uniforms = uniforms;
    // Lognormal of mean `percent` and deviation roughness x percent
    // (lognormal_from_mean_std_device, JuicerCudaScanPipeline.cu:248-255),
    // its parameters in FP64 once per frame.
    const double mean = std::max(0.0, static_cast<double>(sp.glare_percent));
    const double sd = std::max(0.0, static_cast<double>(sp.glare_roughness) * sp.glare_percent);
    const double sigma_sq = std::log1p((sd * sd) / (mean * mean));
    ga.sigma = static_cast<float>(std::sqrt(std::max(0.0, sigma_sq)));
    ga.mu = static_cast<float>(std::log(std::max(1e-12, mean)) - 0.5 * sigma_sq);
    hipError_t e = launch_glare_generate(ga, stream);
    if (e != hipSuccess || weights.empty()) return e;
    const void* gptr = nullptr;
    e = dev.table(hash_bytes(weights.data(), weights.size() * 4, 0x474Cull), weights.data(), weights.size() * 4, &gptr);
    if (e != hipSuccess) return e;
    GrainBlurArgs h{};
    h.in = field;
    h.out = tmp;
    h.weights = static_cast<const float*>(gptr);
    h.radius = static_cast<int>(weights.size() / 2);
    h.width = width;
    h.height = height;
    e = launch_grain_blur_h(h, stream);
    if (e != hipSuccess) return e;
    GrainBlurArgs v = h;
    v.in = tmp;
    v.out = field;
    v.initialize = 1;
    return launch_grain_blur_v(v, stream);
}

FjRenderOutcome render_frame(DeviceState& dev, const FjFrame& frame, const FjSubmission& submission,
                             const FjPreparedHostData& prepared, FjAbortCallback abort_callback, FjErrorBuffer* error,
                             const FjHostPreparation* host) {
    // Scanner LUT and Gaussian kernels of this frame come from the host's
    // callbacks when it supplied them.
    HostPreparationScope host_scope(host);
    Diag d;
    if (!validate_prepared(prepared, d) || !validate_binding(prepared, frame, submission, d))
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, d.text, error);
    if (!validate_frame(frame, d) || !check_supported(prepared, frame, d) ||
        !check_spatial_dir(prepared, frame, dev.max_line, d) || !check_optics_frame(prepared, frame, d) ||
        !check_grain_frame(prepared, frame, d))
        return outcome_diag(d.category, d.text, error);
    ++dev.frame_serial;
    FjRenderOutcome stopped{};
    if (aborted(abort_callback, stopped, error)) return stopped;

    PixelParams pp{};
    bool ok = false;
    hipError_t e = assemble(dev, prepared, pp, d, ok);
    if (e != hipSuccess) return outcome_hip(e, "table upload", error);
    if (!ok) return outcome_diag(d.category, d.text, error);

    const FjRect& w = frame.render_window;
    const FjImage& s = frame.source;
    const FjImage& o = frame.destination;
    const int comps = static_cast<int>(s.components);
    pp.io.src = reinterpret_cast<const char*>(s.address) + (w.y1 - s.bounds.y1) * s.row_bytes +
                static_cast<ptrdiff_t>(w.x1 - s.bounds.x1) * comps * 4;
    pp.io.dst = reinterpret_cast<char*>(o.address) + (w.y1 - o.bounds.y1) * o.row_bytes +
                static_cast<ptrdiff_t>(w.x1 - o.bounds.x1) * comps * 4;
    pp.io.src_row_bytes = s.row_bytes;
    pp.io.dst_row_bytes = o.row_bytes;
    pp.io.width = w.x2 - w.x1;
    pp.io.height = w.y2 - w.y1;
    pp.io.comps = comps;

    const bool has_stream = (frame.flags & FJ_FRAME_STREAM_PRESENT) != 0;
    hipStream_t stream = has_stream ? reinterpret_cast<hipStream_t>(frame.stream) : nullptr;

    // Earlier asynchronous frames may have reported DIR failures.
    unsigned int earlier = dev.drain_errors(false);

    const int slot = dev.error_next;
    dev.error_next = (dev.error_next + 1) % DeviceState::kRing;
    if (dev.error_pending[slot]) {
        e = hipEventSynchronize(dev.error_events[slot]);
        if (e != hipSuccess) return outcome_hip(e, "error ring", error);
        earlier |= dev.error_host[slot];
        dev.error_pending[slot] = false;
    }
    pp.error_word = dev.error_words + slot;
    e = launch_clear_word(pp.error_word, stream);
    if (e != hipSuccess) return outcome_hip(e, "clear error word", error);
    const bool trace = (frame.flags & FJ_FRAME_TRACE_VERBOSE) != 0;
    int stage = 0;
    auto mark = [&](const char* name) {
        // Closes the previous stage (if any) and names the next one.
        if (!trace || stage > DeviceState::kMaxStages) return;
        (void)hipEventRecord(dev.stage_events[stage], stream);
        if (name && stage < DeviceState::kMaxStages) dev.stage_names[stage] = name;
        ++stage;
    };
    const int cus = dev.compute_units;
    const bool spatial_dir = prepared.spatial_dir.hash != 0;
    const bool weave = weave_active(prepared);
    const bool post = post_active(prepared.scanner_post) || weave;
    // The weave resamples the unsharp mask's result, so the mask runs as its
    // own pass first.
    const bool unsharp_apart =
        unsharp_separate(prepared.scanner_post) || (weave && unsharp_active(prepared.scanner_post));
    const bool halation = (prepared.optics.flags & FJ_OPTICS_SCATTER_HALATION) != 0;
    const bool camera_diffusion = (prepared.optics.flags & FJ_OPTICS_CAMERA_DIFFUSION) != 0;
    const bool enlarger_diffusion = pp.print && (prepared.optics.flags & FJ_OPTICS_ENLARGER_DIFFUSION) != 0;
    const bool boost = (prepared.optics.flags & FJ_OPTICS_HIGHLIGHT_BOOST) != 0;
    const bool carrier = halation || camera_diffusion || boost;
    const long long plane = static_cast<long long>(pp.io.width) * pp.io.height;
    if (halation && pp.io.width > dev.halation_max_line)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput halation row above the LDS line limit", error);

    // Diffusion stages: components (host behaviour) and padded sizes.
    const FjDiffusion* diff_desc[2] = {camera_diffusion ? &prepared.optics.camera : nullptr,
                                       enlarger_diffusion ? &prepared.optics.enlarger : nullptr};
    long long padded_floats = 0;
    for (const FjDiffusion* dd : diff_desc) {
        if (!dd) continue;
        double lam[kDiffusionComponents], wts[3][kDiffusionComponents];
        uint32_t category = 0;
        std::string diag;
        if (!stage_components(*dd, lam, wts, category, diag)) return outcome_diag(category, diag, error);
        if (pp.io.width < 2 || pp.io.height < 2)
            return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput diffusion frame side below 2", error);
        int pw = 0, ph = 0;
        diffusion_padded(*dd, pp.io.width, pp.io.height, pw, ph);
        if (pw == 0 || ph == 0 || pw > dev.max_line || ph > dev.max_line)
            return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT,
                                "UnsupportedInput diffusion padded frame side above the LDS line limit", error);
        padded_floats = std::max({padded_floats, 6LL * pw * ph, static_cast<long long>(diffusion_build_floats(*dd, pw, ph))});
    }

    const bool grain = (prepared.grain.flags & FJ_GRAIN_ACTIVE) != 0;
    GrainPlan gp;
    if (grain) {
        bool gok = false, rejected = false;
        std::string gdiag;
        e = grain_prepare(dev, prepared, gp, gok, gdiag, rejected);
        {
            uint32_t category = 0;
            std::string text;
            if (take_host_failure(category, text)) return outcome_diag(category, text, error);
        }
        if (e != hipSuccess) return outcome_hip(e, "grain tables", error);
        if (rejected) return outcome_hip(hipErrorInvalidValue, "print visual grain launch", error);
        if (!gok)
            return outcome_diag(gdiag.rfind("UnsupportedInput", 0) == 0 ? FJ_STATUS_UNSUPPORTED_INPUT
                                                                        : FJ_STATUS_PREPARATION_FAILURE,
                                gdiag, error);
    }

    // Scanner-post Gaussian kernels, built before any launch so that a
    // refused kernel leaves the frame untouched.
    static const float kIdentityTap = 1.0f;
    std::vector<float> lens_w, unsharp_w{kIdentityTap}, glare_w;
    const bool glare = glare_active(prepared.scanner_post);
    if (glare && glare_blurred(prepared.scanner_post)) {
        const FjScannerPostEffects& sp = prepared.scanner_post;
        glare_w = kernel_weights(sp.glare_blur_radius, sp.glare_blur_sigma_px);
        if (glare_w.empty()) return host_or(FJ_STATUS_PREPARATION_FAILURE, "gaussian descriptor is invalid", error);
    }
    if (post) {
        const FjScannerPostEffects& sp = prepared.scanner_post;
        if (lens_active(sp)) {
            lens_w = kernel_weights(sp.lens_blur_radius, sp.lens_blur_sigma_px);
            if (lens_w.empty()) return host_or(FJ_STATUS_PREPARATION_FAILURE, "gaussian descriptor is invalid", error);
        }
        if (unsharp_active(sp)) {
            unsharp_w = kernel_weights(sp.unsharp_radius, sp.unsharp_sigma_px);
            if (unsharp_w.empty()) return host_or(FJ_STATUS_PREPARATION_FAILURE, "gaussian descriptor is invalid", error);
        }
    }

    HalationStages hs;
    int hal_k = 0;
    if (halation) {
        if (!halation_stages(prepared.optics.scatter_halation, pp.io.width, pp.io.height, hs))
            return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT,
                                "UnsupportedInput halation recursion without a stable factored form", error);
        for (int c = 0; c < 3; ++c) hal_k = std::max({hal_k, hs.scatter[c].count, hs.bounce[c].count});
    }
    // Workspace (floats): [DIR 12 planes][carrier 3 + ping-pong 3][halation
    // rows 3K][print-linear 3][post 3, unless it reuses DIR or halation
    // scratch][diffusion padded planes / transfer build scratch].
    long long ws_floats = 0;
    auto take = [&](long long floats) {
        const long long o = ws_floats;
        ws_floats += (floats + 63) / 64 * 64;
        return o;
    };
    const long long dir_off = spatial_dir ? take(12 * plane) : -1;
    const long long carrier_off = carrier ? take(6 * plane) : -1;
    const long long q_off = (halation && hal_k > 0) ? take(3LL * hal_k * plane) : -1;
    const long long plin_off = enlarger_diffusion ? take(3 * plane) : -1;
    const long long dens_off = grain ? take(3 * plane) : -1;
    const long long grain_off = grain ? take(5 * plane) : -1;
    long long post_off = -1;
    if (post) post_off = spatial_dir ? dir_off + 9 * plane : (q_off >= 0 ? q_off : take(3 * plane));
    const long long lens_off =
        (post && (lens_active(prepared.scanner_post) || unsharp_apart)) ? take(plane) : -1;
    const long long pad_off = padded_floats ? take(padded_floats) : -1;
    const long long glare_off = glare ? take(plane) : -1;
    const long long glare_tmp_off = glare_w.empty() ? -1 : take(plane);
    const long long hb_off =
        boost ? take(static_cast<long long>((sizeof(HbState) + sizeof(float) - 1) / sizeof(float))) : -1;
    const bool auto_exposure = (prepared.film_exposure.flags & FJ_FILM_AUTO_EXPOSURE) != 0;
    const long long ae_off =
        auto_exposure ? take(static_cast<long long>((sizeof(AeScratch) + sizeof(float) - 1) / sizeof(float))) : -1;
    float* ws = nullptr;
    // Every return after the slot is acquired records the slot's done event
    // on this frame's stream, so that the slot's next user waits for the
    // launches this frame already queued into it.
    struct SlotUse {
        DeviceState& dev;
        hipStream_t stream;
        bool held = false;
        hipError_t release() {
            if (!held) return hipSuccess;
            held = false;
            return dev.release_workspace(stream);
        }
        ~SlotUse() { (void)release(); }
    } slot_use{dev, stream};
    if (ws_floats) {
        e = dev.ensure_workspace(static_cast<size_t>(ws_floats) * sizeof(float));
        if (e != hipSuccess) return outcome_hip(e, "workspace", error);
        ws = static_cast<float*>(dev.workspace());
        e = dev.acquire_workspace(stream);
        if (e != hipSuccess) return outcome_hip(e, "workspace order", error);
        slot_use.held = true;
    }
    // Preparation done: tables resident, workspace ordered.
    if (aborted(abort_callback, stopped, error)) return stopped;
    if (post) {
        pp.linear_out = 1;
        pp.linear_planes = ws + post_off;
    }
    if (auto_exposure) {
        // Metered before every exposure launch of the frame, into this
        // frame's workspace (concurrent frames meter into their own slots).
        AeScratch* ae = reinterpret_cast<AeScratch*>(ws + ae_off);
        AeArgs aa{};
        aa.src = reinterpret_cast<const unsigned char*>(s.address);
        aa.row_bytes = s.row_bytes;
        aa.comps = comps;
        aa.desc = prepared.auto_exposure;
        aa.ex = pp.ex;
        aa.partials = ae->partials;
        aa.histogram = ae->histogram;
        aa.max_y = &ae->max_y;
        aa.scale = &ae->scale;
        mark("auto_exposure");
        e = launch_auto_exposure(aa, stream);
        if (e != hipSuccess) return outcome_hip(e, "auto exposure", error);
        pp.ex.ae = &ae->scale;
        pp.ex.with_route = 1;
    }
    if (glare) {
        // The scan's glare field, generated and blurred before any stage
        // that scans (launch_focused_scan_linear_rgb,
        // JuicerCudaScanPipeline.cu:2361-2405); seeded by the recipe, the
        // frame token and the source origin (JuicerCudaExecutor.cpp:2416-2420).
        const uint64_t seed_in[4] = {prepared.recipe_hash, submission.frame_token,
                                     static_cast<uint64_t>(static_cast<int64_t>(frame.source.bounds.x1)),
                                     static_cast<uint64_t>(static_cast<int64_t>(frame.source.bounds.y1))};
        mark("scanner_glare");
        e = launch_glare(dev, prepared.scanner_post, glare_w, fnv1a_u64s(seed_in, 4), pp.print ? 1u : 0u,
                         frame.source.bounds.x1, frame.source.bounds.y1, pp.io.width, pp.io.height, ws + glare_off,
                         glare_w.empty() ? nullptr : ws + glare_tmp_off, nullptr, stream);
        if (e != hipSuccess) return outcome_hip(e, "scanner glare", error);
        pp.sc.glare = ws + glare_off;
        for (int c = 0; c < 3; ++c)
            pp.sc.glare_xyz[c] = static_cast<float>(0.01 * static_cast<double>(prepared.output_color.illuminant_xyz[c]));
    }
    // Transfer tables first (built once per PSF and padded size).
    DiffusionStage diff_stage[2];
    for (int k = 0; k < 2; ++k) {
        if (!diff_desc[k]) continue;
        e = diffusion_prepare(dev, *diff_desc[k], pp.io.width, pp.io.height, ws + pad_off, stream, diff_stage[k]);
        if (e == hipErrorInvalidValue)
            return outcome_diag(FJ_STATUS_PREPARATION_FAILURE, "InvalidDiffusionHostBehavior field=component_output", error);
        if (e != hipSuccess) return outcome_hip(e, "diffusion transfer", error);
    }
    if (carrier) {
        for (int c = 0; c < 3; ++c) pp.carrier[c] = ws + carrier_off + c * plane;
        mark("camera_carrier");
        e = launch_pixel_stage(pp, kCarrier, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "camera carrier stage", error);
        if (boost) {
            // expose_print_camera / expose_direct_camera
            // (JuicerCudaExecutor.cpp:1886-1960): boosted right after the
            // exposure, before camera diffusion and halation.
            mark("highlight_boost");
            const FjHighlightBoost& hb = prepared.optics.highlight_boost;
            HbArgs ha{};
            for (int c = 0; c < 3; ++c) ha.planes[c] = pp.carrier[c];
            ha.n = static_cast<long long>(pp.io.width) * pp.io.height;
            ha.protection = hb.protection;
            ha.sharpness = hb.sharpness;
            ha.peak_gain = hb.peak_gain;
            ha.state = reinterpret_cast<HbState*>(ws + hb_off);
            ha.error_word = pp.error_word;
            e = launch_highlight_boost(ha, stream);
            if (e != hipSuccess) return outcome_hip(e, "highlight boost", error);
        }
        // Descriptor channels are R, G, B; carrier planes are B/G/R-sensitive.
        float* cur[3];
        float* alt[3];
        for (int c = 0; c < 3; ++c) {
            cur[c] = pp.carrier[2 - c];
            alt[c] = ws + carrier_off + (3 + 2 - c) * plane;
        }
        if (camera_diffusion) {
            mark("camera_diffusion");
            e = run_diffusion(dev, diff_stage[0], cur, pp.io.width, pp.io.height, ws + pad_off, stream, [&](int step) {
                mark(step == 1 ? "camera_diffusion.filter" : "camera_diffusion.inverse");
            });
            if (e != hipSuccess) return outcome_hip(e, "camera diffusion stage", error);
        }
        if (halation) {
            HalArgs ha{};
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < hal_k; ++k) ha.q[c][k] = ws + q_off + (c * hal_k + k) * plane;
            auto fir_only = [](const HalStageDev* st) {
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < st[c].count; ++k)
                        if (st[c].g[k].kind > kHalFir) return false;
                return true;
            };
            // FIR-only stages run tiled from one buffer into the other; the
            // recursive stages run in place.
            auto run_stage = [&](const HalStageDev* host, const HalStageDev* device) -> hipError_t {
                ha.stage = device;
                if (fir_only(host)) {
                    for (int c = 0; c < 3; ++c) {
                        ha.in[c] = cur[c];
                        ha.out[c] = alt[c];
                    }
                    hipError_t r = launch_halation_fir_stage(ha, stream);
                    for (int c = 0; c < 3; ++c) std::swap(cur[c], alt[c]);
                    return r;
                }
                for (int c = 0; c < 3; ++c) {
                    ha.in[c] = cur[c];
                    ha.out[c] = cur[c];
                }
                ha.column_fir = 0;
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < host[c].count; ++k)
                        if (host[c].g[k].kind == kHalFir) ha.column_fir = 1;
                return launch_halation_stage(ha, stream);
            };
            ha.width = pp.io.width;
            ha.height = pp.io.height;
            ha.row_len = hs.row_len;
            ha.col_len = hs.col_len;
            const void* stage_ptr = nullptr;
            const uint64_t hkey = hash_bytes(&hs, sizeof(hs), 0x4841ull);
            e = dev.table(hkey, &hs, sizeof(hs), &stage_ptr);
            if (e != hipSuccess) return outcome_hip(e, "halation tables", error);
            const char* dstages = static_cast<const char*>(stage_ptr);
            if (hs.scatter_active) {
                mark("halation_scatter");
                e = run_stage(hs.scatter, reinterpret_cast<const HalStageDev*>(dstages + offsetof(HalationStages, scatter)));
                if (e != hipSuccess) return outcome_hip(e, "halation scatter stage", error);
            }
            if (hs.bounce_active) {
                mark("halation_bounce");
                e = run_stage(hs.bounce, reinterpret_cast<const HalStageDev*>(dstages + offsetof(HalationStages, bounce)));
                if (e != hipSuccess) return outcome_hip(e, "halation back-reflection stage", error);
            }
        }
        for (int c = 0; c < 3; ++c) pp.carrier[2 - c] = cur[c];
    }
    if (enlarger_diffusion)
        for (int c = 0; c < 3; ++c) pp.plin[c] = ws + plin_off + c * plane;
    if (grain)
        for (int c = 0; c < 3; ++c) pp.dens[c] = ws + dens_off + c * plane;
    if (spatial_dir) {
        for (int c = 0; c < 6; ++c) pp.dir_planes[c] = ws + dir_off + c * plane;
        mark("dir_source");
        e = launch_pixel_stage(pp, kDirSource, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "DIR source stage", error);
        DctFilterJob job;
        job.width = pp.io.width;
        job.height = pp.io.height;
        job.planes = 3;
        {
            std::string diag;
            bool tok = false;
            e = dir_transfer_cached(dev, prepared.spatial_dir, pp.io.width, pp.io.height, job.transfer, diag, tok);
            if (e != hipSuccess) return outcome_hip(e, "DIR transfer upload", error);
            if (!tok) return host_or(FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=" + diag, error);
        }
        for (int c = 0; c < 3; ++c) {
            job.in[c] = pp.dir_planes[c];
            job.out[c] = pp.dir_planes[c];
            job.tmp_a[c] = ws + dir_off + (6 + c) * plane;
            job.tmp_b[c] = ws + dir_off + (9 + c) * plane;
        }
        mark("dir_dct_filter");
        e = run_dct_filter(dev, job, stream);
        if (e != hipSuccess) return outcome_hip(e, "DIR spatial filter", error);
        mark("pixel_chain");
        e = launch_pixel_stage(pp, kDirFinal, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "DIR final stage", error);
    } else {
        mark("pixel_chain");
        e = launch_pixel_stage(pp, kFused, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "fused pixel kernel", error);
    }
    if (grain) {
        // develop_print_capture (JuicerCudaExecutor.cpp:2198-2280): visual
        // grain on the developed density, then the print exposure.
        if (dev.grain_capture[0]) {
            if (plane <= dev.grain_capture_samples)
                for (int c = 0; c < 3 && e == hipSuccess; ++c)
                    e = hipMemcpyAsync(dev.grain_capture[c], pp.dens[c], static_cast<size_t>(plane) * sizeof(float),
                                       hipMemcpyDeviceToDevice, stream);
            dev.grain_capture[0] = dev.grain_capture[1] = dev.grain_capture[2] = nullptr;
            if (e != hipSuccess) return outcome_hip(e, "grain input capture", error);
        }
        mark("visual_grain");
        e = run_grain(gp, pp.dens, pp.io.width, pp.io.height, ws + grain_off, stream);
        if (e != hipSuccess) return outcome_hip(e, "print visual grain launch", error);
        mark(enlarger_diffusion ? "print_linear" : "print_scan");
        e = launch_pixel_stage(pp, kDensityFinal, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "density final stage", error);
    }
    if (enlarger_diffusion) {
        // develop_print_medium (JuicerCudaExecutor.cpp:2282-2382): enlarger
        // diffusion of the print-linear exposure, then the print and scan.
        mark("enlarger_diffusion");
        e = run_diffusion(dev, diff_stage[1], pp.plin, pp.io.width, pp.io.height, ws + pad_off, stream, [&](int step) {
            mark(step == 1 ? "enlarger_diffusion.filter" : "enlarger_diffusion.inverse");
        });
        if (e != hipSuccess) return outcome_hip(e, "enlarger diffusion stage", error);
        mark("print_scan");
        e = launch_pixel_stage(pp, kEnlargerFinal, cus, stream);
        if (e != hipSuccess) return outcome_hip(e, "enlarger print stage", error);
    }
    if (post) {
        PostParams po{};
        po.io = pp.io;
        po.co = pp.co;
        po.planes = pp.linear_planes;
        const FjScannerPostEffects& sp = prepared.scanner_post;
        const long long np = static_cast<long long>(pp.io.width) * pp.io.height;
        if (lens_active(sp)) {
            const std::vector<float>& lw = lens_w;
            const void* lptr = nullptr;
            e = dev.table(hash_bytes(lw.data(), lw.size() * 4, 0x4C42ull), lw.data(), lw.size() * 4, &lptr);
            if (e != hipSuccess) return outcome_hip(e, "lens blur weights", error);
            mark("post_lens_blur");
            for (int c = 0; c < 3 && e == hipSuccess; ++c) {
                GrainBlurArgs h{};
                h.in = pp.linear_planes + c * np;
                h.out = ws + lens_off;
                h.weights = static_cast<const float*>(lptr);
                h.radius = sp.lens_blur_radius;
                h.width = pp.io.width;
                h.height = pp.io.height;
                e = launch_grain_blur_h(h, stream);
                if (e != hipSuccess) break;
                GrainBlurArgs v = h;
                v.in = ws + lens_off;
                v.out = pp.linear_planes + c * np;
                v.initialize = 1;
                e = launch_grain_blur_v(v, stream);
            }
            if (e != hipSuccess) return outcome_hip(e, "lens blur", error);
        }
        const std::vector<float>& w = unsharp_w;
        const void* wptr = nullptr;
        e = dev.table(hash_bytes(w.data(), w.size() * 4, 0x5057ull), w.data(), w.size() * 4, &wptr);
        if (e != hipSuccess) return outcome_hip(e, "unsharp weights", error);
        const bool fused_unsharp = unsharp_active(sp) && !unsharp_apart;
        if (unsharp_apart) {
            // x + a (x - G * x) per plane: horizontal blur into the scratch
            // plane, then the vertical blur with the unsharp combine in place.
            mark("post_unsharp_wide");
            for (int c = 0; c < 3 && e == hipSuccess; ++c) {
                GrainBlurArgs h{};
                h.in = pp.linear_planes + c * np;
                h.out = ws + lens_off;
                h.weights = static_cast<const float*>(wptr);
                h.radius = sp.unsharp_radius;
                h.width = pp.io.width;
                h.height = pp.io.height;
                e = launch_grain_blur_h(h, stream);
                if (e != hipSuccess) break;
                GrainBlurArgs v = h;
                v.in = ws + lens_off;
                v.out = pp.linear_planes + c * np;
                v.unsharp = 1;
                v.weight = sp.unsharp_amount;
                e = launch_grain_blur_v(v, stream);
            }
            if (e != hipSuccess) return outcome_hip(e, "unsharp mask", error);
        }
        po.weights = static_cast<const float*>(wptr);
        po.radius = fused_unsharp ? sp.unsharp_radius : 0;
        po.amount = fused_unsharp ? sp.unsharp_amount : 0.0f;
        if (weave) {
            po.weave = 1;
            po.weave_dx = prepared.effects.weave_dx_px;
            po.weave_dy = prepared.effects.weave_dy_px;
            po.weave_cos = prepared.effects.weave_cos_rot;
            po.weave_sin = prepared.effects.weave_sin_rot;
        }
        mark("post_unsharp_encode");
        e = launch_post(po, stream);
        if (e != hipSuccess) return outcome_hip(e, "post stage", error);
    }
    mark(nullptr);
    dev.stage_count = trace ? std::min(stage - 1, static_cast<int>(DeviceState::kMaxStages)) : dev.stage_count;
    // Every launch is queued; a cancellation now still lets them run.
    if (aborted(abort_callback, stopped, error)) return stopped;
    e = slot_use.release();
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error);
    e = hipMemcpyAsync(dev.error_host + slot, pp.error_word, sizeof(unsigned int), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return outcome_hip(e, "error readback", error);
    e = hipEventRecord(dev.error_events[slot], stream);
    if (e != hipSuccess) return outcome_hip(e, "error event", error);
    dev.error_pending[slot] = true;
    if (boost) {
        // The boost's disposition is checked before the frame returns
        // (check_highlight_boost, ProcessRoot.cpp:4063-4096).
        e = hipEventSynchronize(dev.error_events[slot]);
        if (e != hipSuccess) return outcome_hip(e, "highlight boost status", error);
        dev.error_pending[slot] = false;
        const unsigned int bits = dev.error_host[slot];
        earlier |= bits;
        if (bits & kHbInvalidCarrierBit)
            return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostInvalidCarrier", error);
        if (bits & kHbPeakOverflowBit)
            return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostPeakOverflow", error);
    }

    FjRenderOutcome out{};
    out.status = status_of(FJ_STATUS_SUCCESS);
    if (!has_stream) {
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return outcome_hip(e, "stream synchronize", error);
        earlier |= dev.drain_errors(false);
    }
    if (earlier & (1u << 8)) out.flags |= FJ_RENDER_DEFERRED_DIR_ERROR;
    return out;
}

FjStatus render_batch(DeviceState& dev, int count, const FjFrame* frames, const FjSubmission* submissions,
                      const FjPreparedHostData* const* prepared, int streams, hipStream_t caller,
                      FjAbortCallback abort_callback, FjRenderOutcome* outcomes, FjErrorBuffer* error) {
    streams = std::max(1, std::min(streams, DeviceState::kWorkspaceSlots));
    hipError_t e = hipSuccess;
    for (int k = 0; k < streams; ++k) {
        if (!dev.workers[k]) {
            e = hipStreamCreateWithFlags(&dev.workers[k], hipStreamNonBlocking);
            if (e != hipSuccess) return outcome_hip(e, "batch streams", error).status;
        }
    }
    // Everything the caller queued before the batch precedes every frame.
    e = hipEventRecord(dev.batch_events[DeviceState::kWorkspaceSlots], caller);
    if (e != hipSuccess) return outcome_hip(e, "batch order", error).status;
    for (int k = 0; k < streams; ++k) {
        e = hipStreamWaitEvent(dev.workers[k], dev.batch_events[DeviceState::kWorkspaceSlots], 0);
        if (e != hipSuccess) return outcome_hip(e, "batch order", error).status;
    }
    FjStatus first = status_of(FJ_STATUS_SUCCESS);
    std::string first_text;
    for (int i = 0; i < count; ++i) {
        const int k = i % streams;
        FjFrame f = frames[i];
        f.stream = reinterpret_cast<uintptr_t>(dev.workers[k]);
        f.flags = (f.flags | FJ_FRAME_STREAM_PRESENT) & ~FJ_FRAME_TRACE_VERBOSE;
        dev.active_slot = k;
        char text[512];
        FjErrorBuffer local{text, sizeof(text), 0};
        outcomes[i] = render_frame(dev, f, submissions[i], *prepared[i], abort_callback, &local);
        if (outcomes[i].status.category != FJ_STATUS_SUCCESS && first.category == FJ_STATUS_SUCCESS) {
            first = outcomes[i].status;
            first_text = "frame " + std::to_string(i) + ": " + std::string(text, local.length);
        }
    }
    dev.active_slot = 0;
    // The caller's stream continues after every worker's last frame.
    for (int k = 0; k < streams; ++k) {
        e = hipEventRecord(dev.batch_events[k], dev.workers[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(caller, dev.batch_events[k], 0);
        if (e != hipSuccess) return outcome_hip(e, "batch join", error).status;
    }
    put_error(error, first_text);
    return first;
}

FjStatus grain_apply_planes(DeviceState& dev, const FjPreparedHostData& prepared, float* const density[3], int width,
                            int height, hipStream_t stream, FjErrorBuffer* error) {
    Diag d;
    if (!validate_prepared(prepared, d)) return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, d.text, error).status;
    std::string field;
    if (!grain_decode(prepared.grain, field)) {
        d.malformed(field.c_str());
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, d.text, error).status;
    }
    if ((prepared.grain.flags & FJ_GRAIN_ACTIVE) == 0 || width <= 0 || height <= 0 || !density[0] || !density[1] ||
        !density[2] || prepared.grain.render_extent.width != width || prepared.grain.render_extent.height != height)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput grain planes do not match an active descriptor",
                            error)
            .status;
    ++dev.frame_serial;
    GrainPlan gp;
    bool ok = false, rejected = false;
    std::string diag;
    hipError_t e = grain_prepare(dev, prepared, gp, ok, diag, rejected);
    if (e != hipSuccess) return outcome_hip(e, "grain tables", error).status;
    if (rejected) return outcome_hip(hipErrorInvalidValue, "visual grain launch", error).status;
    if (!ok)
        return outcome_diag(diag.rfind("UnsupportedInput", 0) == 0 ? FJ_STATUS_UNSUPPORTED_INPUT
                                                                    : FJ_STATUS_PREPARATION_FAILURE,
                            diag, error)
            .status;
    const long long plane = static_cast<long long>(width) * height;
    e = dev.ensure_workspace(static_cast<size_t>(plane) * 5 * sizeof(float));
    if (e != hipSuccess) return outcome_hip(e, "workspace", error).status;
    e = dev.acquire_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    e = run_grain(gp, density, width, height, static_cast<float*>(dev.workspace()), stream);
    if (e != hipSuccess) return outcome_hip(e, "visual grain launch", error).status;
    e = dev.release_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    put_error(error, "");
    return status_of(FJ_STATUS_SUCCESS);
}

FjStatus glare_field(DeviceState& dev, const FjScannerPostEffects& post, uint64_t seed, uint32_t medium, int origin_x,
                     int origin_y, int width, int height, float* field, float* uniforms, hipStream_t stream,
                     FjErrorBuffer* error) {
    if (!glare_active(post) || !post_kernels_valid(post))
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=scanner_post.glare", error).status;
    if (!field || width <= 0 || height <= 0 || medium > 1)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput glare field", error).status;
    std::vector<float> weights;
    if (glare_blurred(post)) {
        weights = kernel_weights(post.glare_blur_radius, post.glare_blur_sigma_px);
        if (weights.empty())
            return outcome_diag(FJ_STATUS_PREPARATION_FAILURE, "gaussian descriptor is invalid", error).status;
    }
    ++dev.frame_serial;
    const size_t plane_bytes = static_cast<size_t>(width) * height * sizeof(float);
    hipError_t e = dev.ensure_workspace(weights.empty() ? sizeof(float) : plane_bytes);
    if (e != hipSuccess) return outcome_hip(e, "workspace", error).status;
    e = dev.acquire_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    e = launch_glare(dev, post, weights, seed, medium, origin_x, origin_y, width, height, field,
                     static_cast<float*>(dev.workspace()), uniforms, stream);
    const hipError_t r = dev.release_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "scanner glare", error).status;
    if (r != hipSuccess) return outcome_hip(r, "workspace order", error).status;
    put_error(error, "");
    return FjStatus{FJ_STATUS_SUCCESS, FJ_API_NONE, 0};
}

FjStatus highlight_boost_planes(DeviceState& dev, const FjHighlightBoost& boost, float* const planes[3],
                                long long samples, hipStream_t stream, FjErrorBuffer* error) {
    if (boost.recipe_hash == 0 || !highlight_boost_valid(boost))
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=optics.highlight_boost", error)
            .status;
    if (samples < 0 || !planes[0] || !planes[1] || !planes[2])
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput highlight boost planes", error).status;
    ++dev.frame_serial;
    hipError_t e = dev.ensure_workspace(sizeof(HbState));
    if (e != hipSuccess) return outcome_hip(e, "workspace", error).status;
    e = dev.acquire_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    const int slot = dev.error_next;
    dev.error_next = (dev.error_next + 1) % DeviceState::kRing;
    if (dev.error_pending[slot]) {
        e = hipEventSynchronize(dev.error_events[slot]);
        if (e != hipSuccess) return outcome_hip(e, "error ring", error).status;
        dev.error_pending[slot] = false;
    }
    HbArgs ha{};
    for (int c = 0; c < 3; ++c) ha.planes[c] = planes[c];
    ha.n = samples;
    ha.protection = boost.protection;
    ha.sharpness = boost.sharpness;
    ha.peak_gain = boost.peak_gain;
    ha.state = static_cast<HbState*>(dev.workspace());
    ha.error_word = dev.error_words + slot;
    e = launch_clear_word(ha.error_word, stream);
    if (e == hipSuccess) e = launch_highlight_boost(ha, stream);
    const hipError_t r = dev.release_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "highlight boost", error).status;
    if (r != hipSuccess) return outcome_hip(r, "workspace order", error).status;
    e = hipMemcpyAsync(dev.error_host + slot, ha.error_word, sizeof(unsigned int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(dev.error_events[slot], stream);
    if (e == hipSuccess) e = hipEventSynchronize(dev.error_events[slot]);
    if (e != hipSuccess) return outcome_hip(e, "highlight boost status", error).status;
    const unsigned int bits = dev.error_host[slot];
    if (bits & kHbInvalidCarrierBit)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostInvalidCarrier", error).status;
    if (bits & kHbPeakOverflowBit) return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostPeakOverflow", error).status;
    put_error(error, "");
    return status_of(FJ_STATUS_SUCCESS);
}

FjStatus diffusion_apply_planes(DeviceState& dev, const FjDiffusion& stage, float* const planes[3], int width,
                                int height, hipStream_t stream, FjErrorBuffer* error) {
    if (stage.family > 3 || stage.hash == 0 || stage.sample_hash == 0)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=optics.diffusion", error).status;
    if (width < 2 || height < 2 || !planes[0] || !planes[1] || !planes[2])
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput diffusion frame side below 2", error).status;
    double lam[kDiffusionComponents], wts[3][kDiffusionComponents];
    uint32_t category = 0;
    std::string diag;
    if (!stage_components(stage, lam, wts, category, diag)) return outcome_diag(category, diag, error).status;
    int pw = 0, ph = 0;
    diffusion_padded(stage, width, height, pw, ph);
    if (pw == 0 || ph == 0 || pw > dev.max_line || ph > dev.max_line)
        return outcome_diag(FJ_STATUS_UNSUPPORTED_INPUT,
                            "UnsupportedInput diffusion padded frame side above the LDS line limit", error)
            .status;
    ++dev.frame_serial;
    const size_t floats =
        std::max(static_cast<size_t>(6) * pw * ph, diffusion_build_floats(stage, pw, ph));
    hipError_t e = dev.ensure_workspace(floats * sizeof(float));
    if (e != hipSuccess) return outcome_hip(e, "workspace", error).status;
    e = dev.acquire_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    float* ws = static_cast<float*>(dev.workspace());
    DiffusionStage st;
    e = diffusion_prepare(dev, stage, width, height, ws, stream, st);
    if (e == hipErrorInvalidValue)
        return outcome_diag(FJ_STATUS_PREPARATION_FAILURE, "InvalidDiffusionHostBehavior field=component_output", error)
            .status;
    if (e != hipSuccess) return outcome_hip(e, "diffusion transfer", error).status;
    e = run_diffusion(dev, st, planes, width, height, ws, stream);
    if (e != hipSuccess) return outcome_hip(e, "diffusion stage", error).status;
    e = dev.release_workspace(stream);
    if (e != hipSuccess) return outcome_hip(e, "workspace order", error).status;
    put_error(error, "");
    return status_of(FJ_STATUS_SUCCESS);
}

}  // namespace fjx
