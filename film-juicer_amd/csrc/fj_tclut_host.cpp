// fj_tclut_host.cpp — the film TC-LUT build: validation, scalar preparation
// and the GPU passes (fj_tclut.hip) behind fjx_build_tc_lut.
//
// Follows build_tc_lut (rust/film-juicer-core/src/reconstruction.rs:594-677)
// step by step, with its failure order and diagnostics: reference
// illuminant; Hanatos spectral blur (kernel :227-268, blur :540-569); Arctic
// neutral response of the projected white (:609-626, Mitchell sampling
// :501-539); Hanatos surface white (:627-637); the integration (:638-657);
// the input-gamut remap (gamut.rs:343-364). The scalar steps that call libm
// (expf, exp2f, sqrtf) run here in FP32 exactly as the reference evaluates
// them; the table passes run on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fj/fj_mi355x_ext.h"
#include "fj_owner.h"
#include "fj_tclut.h"

#pragma clang fp contract(off)

namespace fjx {

namespace {

enum TcFailure {
    kTcOk = 0,
    kTcInvalidIlluminant,
    kTcInvalidBlur,
    kTcInvalidKernel,
    kTcImpossibleKernel,
    kTcInvalidArcticWhite,
    kTcInvalidNeutral,
    kTcInvalidSurfaceWhite,
    kTcNonfiniteIntegrated,
    kTcNonfiniteBlurred,
    kTcMissingHull,
    kTcUnusableCompressed,
    kTcNonfiniteRemapped,
    kTcUnusableSampling,
};

// TcError::diagnostic (reconstruction.rs:441-485).
const char* tc_text(int f) {
    switch (f) {
        case kTcInvalidIlluminant:
            return "MalformedRequiredResource component=film_tc_lut requirement=finite_nonnegative_reference_illuminant";
        case kTcInvalidBlur:
            return "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blur";
        case kTcInvalidKernel:
            return "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blur_kernel";
        case kTcImpossibleKernel:
            return "MalformedRequiredResource component=film_tc_lut requirement=representable_blur_kernel";
        case kTcInvalidArcticWhite:
            return "MalformedRequiredResource component=film_tc_lut method=arctic2026beta04 "
                   "requirement=finite_d65_projection_white";
        case kTcInvalidNeutral:
            return "MalformedRequiredResource component=film_tc_lut method=arctic2026beta04 "
                   "requirement=finite_positive_neutral_response";
        case kTcInvalidSurfaceWhite:
            return "MalformedRequiredResource component=film_tc_lut method=hanatos2025 "
                   "requirement=finite_surface_reference_white";
        case kTcNonfiniteIntegrated:
            return "MalformedRequiredResource component=film_tc_lut requirement=finite_integrated_samples";
        case kTcNonfiniteBlurred:
            return "MalformedRequiredResource component=film_tc_lut method=hanatos2025 requirement=finite_blurred_spectra";
        case kTcMissingHull:
            return "MissingRequiredResource component=film_tc_lut requirement=input_compression_hull";
        case kTcUnusableCompressed:
            return "MalformedRequiredResource component=input_gamut_remap requirement=finite_compressed_coordinates";
        case kTcNonfiniteRemapped:
            return "MalformedRequiredResource component=input_gamut_remap requirement=finite_remapped_tc_lut";
        case kTcUnusableSampling:
            return "MalformedRequiredResource component=film_tc_lut requirement=usable_sampling_coordinates";
        default:
            return "";
    }
}

// The largest blur radius admitted; beyond it the kernel is refused as not
// representable (the reference's bound is the allocator's).
constexpr long kMaxBlurRadius = 1L << 20;

float clamp01(float v) { return std::isnan(v) ? v : (v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v)); }

void tri2quad(float tx, float ty, float out[2]) {
    const float x = 1.0f - tx;
    const float den = x < 1e-10f ? 1e-10f : x;
    out[0] = clamp01(x * x);
    out[1] = clamp01(ty / den);
}

bool projected_white(const float w[3], float out[2]) {
    const float sum = (w[0] + w[1]) + w[2];
    if (!std::isfinite(sum) || sum <= 0.0f) return false;
    tri2quad(w[0] / sum, w[1] / sum, out);
    return std::isfinite(out[0]) && std::isfinite(out[1]);
}

float mitchell_weight(float t) {
    constexpr float B = 1.0f / 3.0f, C = 1.0f / 3.0f;
    const float x = std::fabs(t);
    if (x < 1.0f)
        return (1.0f / 6.0f) * ((12.0f - 9.0f * B - 6.0f * C) * x * x * x + (-18.0f + 12.0f * B + 6.0f * C) * x * x +
                                (6.0f - 2.0f * B));
    if (x < 2.0f)
        return (1.0f / 6.0f) * ((-B - 6.0f * C) * x * x * x + (6.0f * B + 30.0f * C) * x * x +
                                (-12.0f * B - 48.0f * C) * x + (8.0f * B + 24.0f * C));
    return 0.0f;
}

bool mitchell_coordinate(float normalized, int& base, float& frac) {
    const float v = clamp01(normalized) * 191.0f;
    if (!std::isfinite(v)) return false;
    if (v >= 191.0f) {
        base = 190;
        frac = 1.0f;
        return true;
    }
    base = static_cast<int>(std::floor(v));
    frac = v - static_cast<float>(base);
    return true;
}

int mitchell_reflect(int i) { return i < 0 ? -i : (i >= 192 ? 382 - i : i); }

// Mitchell-Netravali (B = C = 1/3) sample of component `component` of a
// 192 x 192 table with `stride` floats per cell.
bool sample_mitchell(const float* s, const float tc[2], int stride, int component, float& out) {
    int cb = 0, mb = 0;
    float cf = 0.0f, mf = 0.0f;
    if (!mitchell_coordinate(tc[0], cb, cf) || !mitchell_coordinate(tc[1], mb, mf)) return false;
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
            value += weight * s[(static_cast<size_t>(c) * 192 + m) * stride + component];
        }
    }
    out = weight_sum != 0.0f ? value / weight_sum : 0.0f;
    return true;
}

float hanatos_surface(const float* p, const float tc[2], const float center[2]) {
    const float x = tc[0] - center[0], y = tc[1] - center[1];
    const float x2 = x * x, y2 = y * y, x3 = x2 * x, y3 = y2 * y;
    const float raw = p[1] * x + p[2] * y + p[3] * x2 + p[4] * y2 + p[5] * x * y + p[6] * x3 + p[7] * y3 +
                      p[8] * x2 * y + p[9] * x * y2 + p[10] * x2 * x2 + p[11] * y2 * y2 + p[12] * x3 * y +
                      p[13] * x2 * y2 + p[14] * x * y3;
    const float normalized = raw / 2.0f;
    return raw / std::sqrt(1.0f + normalized * normalized);
}

int blur_kernel(float sigma, std::vector<float>& k) {
    const float rounded = 4.0f * sigma + 0.5f;
    if (!std::isfinite(rounded) || rounded < 1.0f) return kTcInvalidKernel;
    if (static_cast<double>(rounded) >= 9.2233720368547758e18 || static_cast<long>(rounded) > kMaxBlurRadius)
        return kTcImpossibleKernel;
    const long r = static_cast<long>(rounded);
    k.assign(static_cast<size_t>(2 * r + 1), 0.0f);
    const float s2 = sigma * sigma;
    float sum = 0.0f;
    for (long i = 0; i <= 2 * r; ++i) {
        const long off = i - r;
        k[i] = std::exp(-0.5f * static_cast<float>(off * off) / s2);
        sum += k[i];
    }
    if (!std::isfinite(sum) || sum <= 0.0f) return kTcInvalidKernel;
    for (float& v : k) v /= sum;
    return kTcOk;
}

bool span_ok(const FjFloatSpan& s, size_t count) { return s.data != nullptr && s.count == count; }

// Device buffers of one build inside the device's workspace.
struct TcBuffers {
    float* spectra;
    float* blurred;
    float* integrated;
    float* table;
    float* small;  // kernel weights, surface factors, sensitivity, illuminant, hull, grids
    unsigned int* failures;  // blur, integrate, remap
};

}  // namespace

FjStatus build_tc_lut(DeviceState& dev, const FjFilmTcLutInput& in, float* samples, size_t count, hipStream_t stream,
                      FjErrorBuffer* error) {
    auto refuse = [&](const char* field) {
        return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, std::string("MalformedPreparedHostData field=tc_lut.") + field);
    };
    auto tc_fail = [&](int f) { return fail(error, FJ_STATUS_PREPARATION_FAILURE, tc_text(f)); };
    const bool arctic = in.method == 2;
    if (in.method != 0 && in.method != 2) return refuse("method");
    if (in.apply_surface > 1 || in.compression_active > 1 || in.hull_available > 1) return refuse("flags");
    if (!span_ok(in.spectra, static_cast<size_t>(kTcCells) * kTcSamples)) return refuse("spectra");
    if (!span_ok(in.sensitivity_rgb, kTcSamples * 3)) return refuse("sensitivity_rgb");
    if (!span_ok(in.reference_illuminant, kTcSamples)) return refuse("reference_illuminant");
    const bool surface = !arctic && in.apply_surface;
    if (surface && !span_ok(in.surface_rgb, 45)) return refuse("surface_rgb");
    const bool remap = in.compression_active != 0;
    if (remap && in.hull_available && !span_ok(in.hull_xy, 2 * 1025)) return refuse("hull_xy");
    if (!samples || count < static_cast<size_t>(kTcCells) * 4) return refuse("samples");

    const float* illum = in.reference_illuminant.data;
    for (int i = 0; i < kTcSamples; ++i)
        if (!std::isfinite(illum[i]) || !(illum[i] >= 0.0f)) return tc_fail(kTcInvalidIlluminant);
    std::vector<float> kernel;
    const bool blur = !arctic && in.spectral_blur > 0.0f;
    if (blur) {
        if (!std::isfinite(in.spectral_blur) || in.spectral_blur < 0.0f) return tc_fail(kTcInvalidBlur);
        const int f = blur_kernel(in.spectral_blur, kernel);
        if (f) return tc_fail(f);
    }
    float denom[3] = {1.0f, 1.0f, 1.0f};
    if (arctic) {
        float tc[2];
        if (!projected_white(in.projection_white_xyz, tc)) return tc_fail(kTcInvalidArcticWhite);
        denom[0] = denom[1] = denom[2] = 0.0f;
        for (int s = 0; s < kTcSamples; ++s) {
            float reflectance = 0.0f;
            if (!sample_mitchell(in.spectra.data, tc, kTcSamples, s, reflectance)) return tc_fail(kTcUnusableSampling);
            const float relit = reflectance * illum[s];
            for (int c = 0; c < 3; ++c) denom[c] += relit * in.sensitivity_rgb.data[s * 3 + c];
        }
        for (float d : denom)
            if (!std::isfinite(d) || !(d > 0.0f)) return tc_fail(kTcInvalidNeutral);
    }
    std::vector<float> factors;
    if (surface) {
        float center[2];
        if (!projected_white(in.projection_white_xyz, center)) return tc_fail(kTcInvalidSurfaceWhite);
        factors.resize(static_cast<size_t>(kTcCells) * 3);
        for (int c = 0; c < kTcSide; ++c)
            for (int m = 0; m < kTcSide; ++m) {
                const float tc[2] = {static_cast<float>(c) / 191.0f, static_cast<float>(m) / 191.0f};
                for (int ch = 0; ch < 3; ++ch)
                    factors[(static_cast<size_t>(c) * kTcSide + m) * 3 + ch] =
                        std::exp2(hanatos_surface(in.surface_rgb.data + 15 * ch, tc, center));
            }
    }
    std::vector<float> grids(2 * kTcSide);
    for (int i = 0; i < kTcSide; ++i) {
        grids[i] = std::sqrt(static_cast<float>(i) / 191.0f);
        grids[kTcSide + i] = static_cast<float>(i) / 191.0f;
    }

    // Device layout in the workspace: spectra, blurred spectra, integrated
    // table, remapped table, then the small arrays and the failure words.
    const size_t n_spec = static_cast<size_t>(kTcCells) * kTcSamples, n_tab = static_cast<size_t>(kTcCells) * 4;
    const size_t n_small = kernel.size() + factors.size() + kTcSamples * 3 + kTcSamples + 2 * 1025 + grids.size();
    const size_t floats = 2 * n_spec + 2 * n_tab + n_small + 64;
    ++dev.frame_serial;
    hipError_t e = dev.ensure_workspace(floats * sizeof(float));
    if (e != hipSuccess) return hip_fail(error, e, "workspace");
    e = dev.acquire_workspace(stream);
    if (e != hipSuccess) return hip_fail(error, e, "workspace order");
    float* ws = static_cast<float*>(dev.workspace());
    TcBuffers b{};
    b.spectra = ws;
    b.blurred = b.spectra + n_spec;
    b.integrated = b.blurred + n_spec;
    b.table = b.integrated + n_tab;
    b.small = b.table + n_tab;
    b.failures = reinterpret_cast<unsigned int*>(b.small + n_small);
    float* d_kernel = b.small;
    float* d_factors = d_kernel + kernel.size();
    float* d_sens = d_factors + factors.size();
    float* d_ill = d_sens + kTcSamples * 3;
    float* d_hull = d_ill + kTcSamples;
    float* d_grids = d_hull + 2 * 1025;
    const unsigned int none[3] = {kTcNoFailure, kTcNoFailure, kTcNoFailure};
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    };
    up(b.spectra, in.spectra.data, n_spec * sizeof(float));
    up(d_kernel, kernel.data(), kernel.size() * sizeof(float));
    up(d_factors, factors.data(), factors.size() * sizeof(float));
    up(d_sens, in.sensitivity_rgb.data, kTcSamples * 3 * sizeof(float));
    up(d_ill, illum, kTcSamples * sizeof(float));
    if (remap && in.hull_available) up(d_hull, in.hull_xy.data, 2 * 1025 * sizeof(float));
    up(d_grids, grids.data(), grids.size() * sizeof(float));
    up(b.failures, none, sizeof(none));
    if (e == hipSuccess && blur) {
        TcBlurArgs ba{b.spectra, b.blurred, d_kernel, static_cast<int>(kernel.size() / 2), b.failures};
        e = launch_tc_blur(ba, stream);
    }
    if (e == hipSuccess) {
        TcIntegrateArgs ia{};
        ia.spectra = blur ? b.blurred : b.spectra;
        ia.sensitivity = d_sens;
        ia.illuminant = d_ill;
        ia.relight = arctic ? 1 : 0;
        for (int c = 0; c < 3; ++c) ia.denominator[c] = denom[c];
        ia.surface_factor = surface ? d_factors : nullptr;
        ia.out = b.integrated;
        ia.first_failure = b.failures + 1;
        e = launch_tc_integrate(ia, stream);
    }
    const bool run_remap = remap && in.hull_available;
    if (e == hipSuccess && run_remap) {
        TcRemapArgs ra{};
        ra.integrated = b.integrated;
        ra.hull_xy = d_hull;
        ra.hull_center[0] = in.hull_center_xy[0];
        ra.hull_center[1] = in.hull_center_xy[1];
        ra.roots = d_grids;
        ra.steps = d_grids + kTcSide;
        ra.out = b.table;
        ra.first_failure = b.failures + 2;
        e = launch_tc_remap(ra, stream);
    }
    unsigned int failures[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(failures, b.failures, sizeof(failures), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(samples, run_remap ? b.table : b.integrated, n_tab * sizeof(float), hipMemcpyDeviceToHost,
                           stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    const hipError_t r = dev.release_workspace(stream);
    if (e != hipSuccess) return hip_fail(error, e, "TC LUT build");
    if (r != hipSuccess) return hip_fail(error, r, "workspace order");
    auto clear = [&]() { std::memset(samples, 0, n_tab * sizeof(float)); };
    if (failures[0] != kTcNoFailure) return clear(), tc_fail(kTcNonfiniteBlurred);
    if (failures[1] != kTcNoFailure) return clear(), tc_fail(kTcNonfiniteIntegrated);
    if (remap && !in.hull_available) return clear(), tc_fail(kTcMissingHull);
    if (failures[2] != kTcNoFailure) {
        const unsigned int rank = failures[2] & 7u;
        clear();
        return tc_fail(rank == 0 ? kTcUnusableCompressed : rank == 1 ? kTcUnusableSampling : kTcNonfiniteRemapped);
    }
    write_error(error, "");
    return make_status(FJ_STATUS_SUCCESS);
}

}  // namespace fjx
