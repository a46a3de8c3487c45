// fj_print_prep.cpp — host preparation of the print stage: the filtered
// enlarger illuminant, the preflash integral and the print balance.
//
// These are 81-sample computations that the reference's Rust host runs once
// per parameter edit (rust/film-juicer-core/src/print.rs:435-658); they give
// FjPrint its main illuminant, preflash raw CMY and normalizer. They run here
// on the host too, in the reference's arithmetic and operation order (FP32
// filters, FP64 integrals, no contraction), so their words equal the
// reference's. The entry points keep the reference's checked test ABI
// (rust/film-juicer-plugin/src/print_bridge.rs:187-640): output and failure
// records cleared first, raw-record refusals, failure reason plus first
// failing sample or channel, and the one-line diagnostic.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "../../include/fj/fj_mi355x_ext.h"

#pragma clang fp contract(off)

#define FJ_EXPORT __attribute__((visibility("default")))

namespace {

// ---- erf as the reference computes it --------------------------------------
// print.rs calls libm::erff from the Rust libm crate pinned at =0.2.15
// (film-juicer-core/Cargo.toml:12), whose erff and expf are the musl ports of
// FreeBSD msun s_erff.c / e_expf.c (Sun fdlibm): rational approximations on
// [0, 0.84375), [0.84375, 1.25), [1.25, 1/0.35), [1/0.35, 6) and
// erfc(x) = exp(-z^2 - 0.5625) exp((z - x)(z + x) + R/S) / x with z = x
// truncated to 11 mantissa bits. glibc's erff (another fdlibm descendant with a
// different expf and truncation) differs from it in a few words, and the
// reference's fixtures hold the libm words, so it is restated here.

float word_float(uint32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

uint32_t float_word(float f) {
    uint32_t w;
    std::memcpy(&w, &f, 4);
    return w;
}

// expf of FreeBSD e_expf.c: k = round(x / ln2), r = x - k ln2 in hi/lo parts,
// exp(r) = 1 + r + r c / (2 - c) with c = r - r^2 (P1 + r^2 P2), scaled by 2^k.
float fdlibm_expf(float x) {
    constexpr float kLn2Hi = 6.9314575195e-01f, kLn2Lo = 1.4286067653e-06f, kInvLn2 = 1.4426950216e+00f;
    constexpr float kP1 = 1.6666625440e-1f, kP2 = -2.7667332906e-3f;
    uint32_t hx = float_word(x);
    const int sign = static_cast<int>(hx >> 31);
    hx &= 0x7fffffffu;
    if (hx >= 0x42aeac50u) {            // |x| >= 87.33655 or NaN
        if (hx > 0x7f800000u) return x;  // NaN
        if (hx >= 0x42b17218u && !sign) return x * word_float(0x7f000000u);  // overflow
        if (sign && hx >= 0x42cff1b5u) return 0.0f;                          // underflow
    }
    int k;
    float hi, lo;
    if (hx > 0x3eb17218u) {  // |x| > 0.5 ln2
        if (hx > 0x3f851592u)
            k = static_cast<int>(kInvLn2 * x + (sign ? -0.5f : 0.5f));
        else
            k = 1 - sign - sign;
        const float kf = static_cast<float>(k);
        hi = x - kf * kLn2Hi;
        lo = kf * kLn2Lo;
        x = hi - lo;
    } else if (hx > 0x39000000u) {  // |x| > 2^-14
        k = 0;
        hi = x;
        lo = 0.0f;
    } else {
        return 1.0f + x;
    }
    const float xx = x * x;
    const float c = x - xx * (kP1 + xx * kP2);
    const float y = 1.0f + (x * c / (2.0f - c) - lo + hi);
    return k == 0 ? y : std::ldexp(y, k);
}

constexpr float kErx = 8.4506291151e-01f, kEfx8 = 1.0270333290e+00f;
constexpr float kPp[5] = {1.2837916613e-01f, -3.2504209876e-01f, -2.8481749818e-02f, -5.7702702470e-03f,
                          -2.3763017452e-05f};
constexpr float kQq[6] = {0.0f, 3.9791721106e-01f, 6.5022252500e-02f, 5.0813062117e-03f, 1.3249473704e-04f,
                          -3.9602282413e-06f};
constexpr float kPa[7] = {-2.3621185683e-03f, 4.1485610604e-01f, -3.7220788002e-01f, 3.1834661961e-01f,
                          -1.1089469492e-01f, 3.5478305072e-02f, -2.1663755178e-03f};
constexpr float kQa[7] = {0.0f, 1.0642088205e-01f, 5.4039794207e-01f, 7.1828655899e-02f,
                          1.2617121637e-01f, 1.3637083583e-02f, 1.1984500103e-02f};
constexpr float kRa[8] = {-9.8649440333e-03f, -6.9385856390e-01f, -1.0558626175e+01f, -6.2375331879e+01f,
                          -1.6239666748e+02f, -1.8460508728e+02f, -8.1287437439e+01f, -9.8143291473e+00f};
constexpr float kSa[9] = {0.0f, 1.9651271820e+01f, 1.3765776062e+02f, 4.3456588745e+02f, 6.4538726807e+02f,
                          4.2900814819e+02f, 1.0863500214e+02f, 6.5702495575e+00f, -6.0424413532e-02f};
constexpr float kRb[7] = {-9.8649431020e-03f, -7.9928326607e-01f, -1.7757955551e+01f, -1.6063638306e+02f,
                          -6.3756646729e+02f, -1.0250950928e+03f, -4.8351919556e+02f};
constexpr float kSb[8] = {0.0f, 3.0338060379e+01f, 3.2579251099e+02f, 1.5367296143e+03f, 3.1998581543e+03f,
                          2.5530502930e+03f, 4.7452853394e+02f, -2.2440952301e+01f};

// Horner from the highest coefficient down: c[0] + s (c[1] + s (... + s c[n-1])).
template <int N>
float horner(const float (&c)[N], float s, int from = 0) {
    float v = c[N - 1];
    for (int i = N - 2; i >= from; --i) v = c[i] + s * v;
    return v;
}

// 1 + s (c[1] + s (c[2] + ...)), the denominators' form.
template <int N>
float one_plus(const float (&c)[N], float s) { return 1.0f + s * horner(c, s, 1); }

float erfc_tail(uint32_t ix, float x) {
    if (ix < 0x3fa00000u) {  // |x| < 1.25
        const float s = std::fabs(x) - 1.0f;
        const float p = horner(kPa, s), q = one_plus(kQa, s);
        return 1.0f - kErx - p / q;
    }
    x = std::fabs(x);
    const float s = 1.0f / (x * x);
    float r, q;
    if (ix < 0x4036db6du) {  // |x| < 1/0.35
        r = horner(kRa, s);
        q = one_plus(kSa, s);
    } else {
        r = horner(kRb, s);
        q = one_plus(kSb, s);
    }
    const float z = word_float(float_word(x) & 0xffffe000u);
    return fdlibm_expf(-z * z - 0.5625f) * fdlibm_expf((z - x) * (z + x) + r / q) / x;
}

float libm_erff(float x) {
    uint32_t ix = float_word(x);
    const bool negative = (ix >> 31) != 0;
    ix &= 0x7fffffffu;
    if (ix >= 0x7f800000u) return (negative ? -1.0f : 1.0f) + 1.0f / x;  // NaN, +-1 at infinity
    if (ix < 0x3f580000u) {                                              // |x| < 0.84375
        if (ix < 0x31800000u) return 0.125f * (8.0f * x + kEfx8 * x);    // |x| < 2^-28
        const float z = x * x;
        const float r = horner(kPp, z), s = one_plus(kQq, z);
        return x + x * (r / s);
    }
    const float y = ix < 0x40c00000u ? 1.0f - erfc_tail(ix, x) : 1.0f - word_float(0x03800000u);
    return negative ? -y : y;
}

// erfcf of the same library: the erf polynomial below 0.84375 (as 1 - erf, or
// 0.5 - (x - 0.5 + x y) above 1/4 to keep the small result exact), the erfc
// tails up to 28, and 0 or 2 beyond.
float libm_erfcf(float x) {
    uint32_t ix = float_word(x);
    const bool negative = (ix >> 31) != 0;
    ix &= 0x7fffffffu;
    if (ix >= 0x7f800000u) return (negative ? 2.0f : 0.0f) + 1.0f / x;  // NaN, 0 at +inf, 2 at -inf
    if (ix < 0x3f580000u) {                                              // |x| < 0.84375
        if (ix < 0x23800000u) return 1.0f - x;                           // |x| < 2^-56
        const float z = x * x;
        const float r = horner(kPp, z), s = one_plus(kQq, z);
        const float y = r / s;
        if (negative || ix < 0x3e800000u) return 1.0f - (x + x * y);  // x < 1/4
        return 0.5f - (x - 0.5f + x * y);
    }
    if (ix < 0x41e00000u) return negative ? 2.0f - erfc_tail(ix, x) : erfc_tail(ix, x);  // |x| < 28
    const float tiny = word_float(0x03800000u);
    return negative ? 2.0f - tiny : tiny * tiny;
}

// ---- the checked ABI ---------------------------------------------------------

FjStatus status(uint32_t category) { return FjStatus{category, FJ_API_NONE, 0}; }

template <typename T>
bool valid_extent(const T* p, size_t count) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    return p && a % alignof(T) == 0 && count <= static_cast<size_t>(INTPTR_MAX) / sizeof(T) &&
           a + count * sizeof(T) >= a;
}

// The optional diagnostic record: valid when absent, or when its positive
// capacity comes with storage; admission clears it.
bool admit_diagnostic(FjErrorBuffer* e) {
    if (!e) return true;
    if (!valid_extent(e, 1)) return false;
    e->length = 0;
    if (e->capacity == 0) return true;
    if (!e->data) return false;
    e->data[0] = '\0';
    return true;
}

void write_diagnostic(FjErrorBuffer* e, const char* kind, const char* index_name, uint32_t reason, uint32_t index) {
    if (!e) return;
    char text[96];
    const int n = std::snprintf(text, sizeof text, "print %s reason=%u %s=%u", kind, reason, index_name, index);
    if (e->capacity == 0 || n <= 0) return;
    size_t count = static_cast<size_t>(n);
    if (count > e->capacity - 1) count = e->capacity - 1;
    std::memcpy(e->data, text, count);
    e->data[count] = '\0';
    e->length = count;
}

bool span_exact(const FjFloatSpan& s, size_t n) { return s.count == n && valid_extent(s.data, n); }

constexpr uint32_t kNoIndex = 0xFFFFFFFFu;
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;

void fnv_bytes(uint64_t& h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= kFnvPrime;
    }
}

// hash::finite_f32s (rust film-juicer-core/src/hash.rs): FNV-1a over the
// little-endian words with either zero hashed as +0; 0 when one is not finite.
uint64_t finite_f32s(const float* v, size_t n) {
    uint64_t h = kFnvOffset;
    for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(v[i])) return 0;
        const uint32_t w = v[i] == 0.0f ? 0u : float_word(v[i]);
        fnv_bytes(h, &w, 4);
    }
    return h;
}

// ---- the computations ----------------------------------------------------------

// Illuminant::new (print.rs:591-639). Returns the reason; *sample names the
// first failing sample.
uint32_t filtered_illuminant(const FjPrintIlluminantParams& p, const float* source, float* samples,
                             uint32_t* sample) {
    float filters[81][3];
    for (int s = 0; s < 81; ++s) {
        const float wavelength = 380.0f + 5.0f * static_cast<float>(s);
        const float y = 0.5f * libm_erff((wavelength - p.edges_nm[0]) / p.transitions_nm[0]) + 0.5f;
        const float m = wavelength <= 550.0f ? -libm_erff((wavelength - p.edges_nm[1]) / p.transitions_nm[1])
                                             : libm_erff((wavelength - p.edges_nm[2]) / p.transitions_nm[2]);
        const float c = -0.5f * libm_erff((wavelength - p.edges_nm[3]) / p.transitions_nm[3]) + 0.5f;
        filters[s][0] = c;
        filters[s][1] = 0.5f * m + 0.5f;
        filters[s][2] = y;
    }
    float transmittance[3];
    for (int ch = 0; ch < 3; ++ch) transmittance[ch] = std::pow(10.0f, -p.cmy_cc[ch] / 100.0f);
    float energy = 0.0f;
    for (int s = 0; s < 81; ++s) {
        float total = source[s];
        for (int ch = 0; ch < 3; ++ch) total *= 1.0f - (1.0f - filters[s][ch]) * (1.0f - transmittance[ch]);
        if (!std::isfinite(total) || total < 0.0f) {
            *sample = static_cast<uint32_t>(s);
            return FJ_PRINT_ILLUMINANT_REASON_SAMPLE;
        }
        samples[s] = total;
        energy += total;
    }
    if (!std::isfinite(energy) || energy <= 0.0f) return FJ_PRINT_ILLUMINANT_REASON_ENERGY;
    return FJ_PRINT_ILLUMINANT_REASON_NONE;
}

// PreflashRaw::new (print.rs:538-571). Returns the first failing channel or -1.
int preflash_raw(const float* base, const float* light, const float* sensitivity, float cmy[3]) {
    double raw[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < 81; ++s) {
        const double transmitted = std::pow(10.0, -static_cast<double>(base[s])) * static_cast<double>(light[s]);
        const double l = std::isnan(transmitted) ? 0.0 : transmitted;
        for (int ch = 0; ch < 3; ++ch) {
            const float v = sensitivity[s * 3 + ch];
            if (std::isfinite(v)) raw[ch] += l * static_cast<double>(v);
        }
    }
    for (int ch = 0; ch < 3; ++ch)
        if (!std::isfinite(raw[ch]) || raw[ch] < 0.0) return ch;
    // Finite totals may narrow to infinity; the payload admits them later.
    for (int ch = 0; ch < 3; ++ch) cmy[ch] = static_cast<float>(raw[ch]);
    return -1;
}

// One raw RGB integral of Balance::new (print.rs:442-469). Returns the first
// failing channel or -1.
int balance_raw(const FjPrintBalanceInput& in, const float density_cmy[3], float raw[3]) {
    const float* base = in.film_base_density.data;
    const float* dye = in.film_channel_density_cmy.data;
    const float* sensitivity = in.print_sensitivity_rgb.data;
    const float* light = in.filtered_main_illuminant.data;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < 81; ++s) {
        double density = base[s];
        for (int ch = 0; ch < 3; ++ch)
            density += static_cast<double>(density_cmy[ch]) * static_cast<double>(dye[s * 3 + ch]);
        const double l = std::pow(10.0, -density) * static_cast<double>(light[s]);
        if (std::isnan(l)) continue;
        for (int ch = 0; ch < 3; ++ch) {
            const float v = sensitivity[s * 3 + ch];
            if (std::isfinite(v)) acc[ch] += l * static_cast<double>(v);
        }
    }
    for (int ch = 0; ch < 3; ++ch) {
        // The narrowed value is admitted, a negative total narrowing to -0 included.
        raw[ch] = static_cast<float>(acc[ch]);
        if (!std::isfinite(raw[ch]) || raw[ch] < 0.0f) return ch;
    }
    return -1;
}

float midgray_factor(const float raw[3]) {
    double mean_log = 0.0;
    for (int ch = 0; ch < 3; ++ch) mean_log += std::log(std::fmax(static_cast<double>(raw[ch]), 1.0e-10));
    return static_cast<float>(std::exp(-mean_log / 3.0));
}

// Balance::new after the integrals (print.rs:476-521). False for an invalid balance.
bool balance_scalars(const FjPrintBalanceInput& in, FjPrintBalance& b) {
    b.factor_midgray = midgray_factor(b.baseline_raw_rgb);
    b.factor_midgray_comp = midgray_factor(b.compensated_raw_rgb);
    switch (in.normalization_mode) {
        case 0: b.normalizer = 1.0f; break;
        case 1: b.normalizer = b.factor_midgray_comp / b.factor_midgray; break;
        case 2: b.normalizer = b.factor_midgray; break;
        case 3: b.normalizer = b.factor_midgray_comp; break;
        default: b.normalizer = std::numeric_limits<float>::quiet_NaN(); break;
    }
    b.filtered_main_illuminant_hash = finite_f32s(in.filtered_main_illuminant.data, 81);
    const uint64_t lanes[7] = {in.reference_hash,
                               b.filtered_main_illuminant_hash,
                               finite_f32s(b.baseline_raw_rgb, 3),
                               finite_f32s(b.compensated_raw_rgb, 3),
                               float_word(b.factor_midgray),
                               float_word(b.factor_midgray_comp),
                               float_word(b.normalizer)};
    uint64_t h = kFnvOffset;
    for (uint64_t lane : lanes) fnv_bytes(h, &lane, 8);
    b.hash = h;
    return std::isfinite(b.factor_midgray) && b.factor_midgray > 0.0f && std::isfinite(b.factor_midgray_comp) &&
           b.factor_midgray_comp > 0.0f && std::isfinite(b.normalizer) && b.normalizer > 0.0f &&
           b.filtered_main_illuminant_hash != 0 && b.hash != 0;
}

// print_gamma_model (profile/sampled.rs:212-236): centres and sigmas divided by
// the gamma in FP64, sigmas floored at 0.05 by a plain comparison (a NaN sigma
// stays NaN); a gamma of exactly 1 leaves every bit as it is.
FjDensityCurveModel print_gamma_model(const FjDensityCurveModel& m, double gamma) {
    FjDensityCurveModel adjusted = m;
    if (gamma != 1.0) {
        for (int ch = 0; ch < 3; ++ch)
            for (int layer = 0; layer < 3; ++layer) {
                adjusted.centers[ch][layer] /= gamma;
                const double sigma = adjusted.sigmas[ch][layer] / gamma;
                adjusted.sigmas[ch][layer] = sigma < 0.05 ? 0.05 : sigma;
            }
    }
    return adjusted;
}

struct DensityFault {
    uint32_t reason = FJ_PRINT_DENSITY_REASON_NONE, channel = kNoIndex, layer = kNoIndex;
};

// DensityCurveModel::sample (profile.rs:482-524): FP32 from the narrowed
// exposure and coefficients, amplitude x 0.5 erfc(-z / sqrt 2) per layer,
// totals summed in layer order. Writes total[3] and layers[layer][channel].
DensityFault sample_density(const FjDensityCurveModel& m, bool positive, double log_exposure, float total_out[3],
                            float layers_out[3][3]) {
    constexpr float kInverseSqrtTwo = 0.7071067811865475244f;
    const float exposure = static_cast<float>(log_exposure);
    const float sign = positive ? -1.0f : 1.0f;
    for (int ch = 0; ch < 3; ++ch) {
        float total = 0.0f;
        for (int layer = 0; layer < 3; ++layer) {
            const float center = static_cast<float>(m.centers[ch][layer]);
            const float amplitude = static_cast<float>(m.amplitudes[ch][layer]);
            const float sigma = static_cast<float>(m.sigmas[ch][layer]);
            const float z = sign * (exposure - center) / sigma;
            const float value = amplitude * (0.5f * libm_erfcf(-z * kInverseSqrtTwo));
            if (!std::isfinite(value)) {
                DensityFault f;
                f.reason = FJ_PRINT_DENSITY_REASON_NONFINITE_LAYER;
                f.channel = static_cast<uint32_t>(ch);
                f.layer = static_cast<uint32_t>(layer);
                return f;
            }
            layers_out[layer][ch] = value;
            total += value;
        }
        if (!std::isfinite(total)) {
            DensityFault f;
            f.reason = FJ_PRINT_DENSITY_REASON_NONFINITE_TOTAL;
            f.channel = static_cast<uint32_t>(ch);
            return f;
        }
        total_out[ch] = total;
    }
    return DensityFault{};
}

}  // namespace

// The enlarger's dichroic filter curves (C, M, Y per wavelength) of the
// reference's filter model (print.rs:6-8 edges and transitions, :616-633).
namespace fjx {
void dichroic_filters(float filters[81][3]) {
    static const float kEdges[4] = {516.0f, 500.0f, 610.0f, 607.0f};
    static const float kTransitions[4] = {12.0f, 8.0f, 8.0f, 8.0f};
    for (int s = 0; s < 81; ++s) {
        const float wavelength = 380.0f + 5.0f * static_cast<float>(s);
        const float y = 0.5f * libm_erff((wavelength - kEdges[0]) / kTransitions[0]) + 0.5f;
        const float m = wavelength <= 550.0f ? -libm_erff((wavelength - kEdges[1]) / kTransitions[1])
                                             : libm_erff((wavelength - kEdges[2]) / kTransitions[2]);
        const float c = -0.5f * libm_erff((wavelength - kEdges[3]) / kTransitions[3]) + 0.5f;
        filters[s][0] = c;
        filters[s][1] = 0.5f * m + 0.5f;
        filters[s][2] = y;
    }
}
}  // namespace fjx

extern "C" {

FJ_EXPORT FjStatus fjx_prepare_print_illuminant(const FjPrintIlluminantParams* params, FjFloatSpan source,
                                                FjIlluminant* out, FjPrintIlluminantFailure* failure,
                                                FjErrorBuffer* error) {
    const bool valid_out = valid_extent(out, 1), valid_failure = valid_extent(failure, 1);
    if (valid_out) *out = FjIlluminant{};
    if (valid_failure) *failure = FjPrintIlluminantFailure{FJ_PRINT_ILLUMINANT_REASON_NONE, kNoIndex};
    const bool diagnostic_valid = admit_diagnostic(error);
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT, reason = FJ_PRINT_ILLUMINANT_REASON_RAW, sample = kNoIndex;
    if (valid_out && valid_failure && diagnostic_valid && valid_extent(params, 1) && span_exact(source, 81)) {
        float samples[81] = {};
        reason = filtered_illuminant(*params, source.data, samples, &sample);
        if (reason == FJ_PRINT_ILLUMINANT_REASON_NONE) {
            std::memcpy(out->samples, samples, sizeof samples);
            return status(FJ_STATUS_SUCCESS);
        }
        category = FJ_STATUS_PREPARATION_FAILURE;
    }
    if (diagnostic_valid) write_diagnostic(error, "illuminant", "sample", reason, sample);
    if (valid_failure) *failure = FjPrintIlluminantFailure{reason, sample};
    return status(category);
}

FJ_EXPORT FjStatus fjx_prepare_print_preflash(const FjPrintPreflashInput* input, FjPrintPreflashRaw* out,
                                              FjPrintPreflashFailure* failure, FjErrorBuffer* error) {
    const bool valid_out = valid_extent(out, 1), valid_failure = valid_extent(failure, 1);
    if (valid_out) *out = FjPrintPreflashRaw{};
    if (valid_failure) *failure = FjPrintPreflashFailure{FJ_PRINT_PREFLASH_REASON_NONE, kNoIndex};
    const bool diagnostic_valid = admit_diagnostic(error);
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT, reason = FJ_PRINT_PREFLASH_REASON_RAW, channel = kNoIndex;
    if (valid_out && valid_failure && diagnostic_valid && valid_extent(input, 1) &&
        span_exact(input->film_base_density, 81) && span_exact(input->filtered_illuminant, 81) &&
        span_exact(input->print_sensitivity_cmy, 243)) {
        float cmy[3];
        const int failed = preflash_raw(input->film_base_density.data, input->filtered_illuminant.data,
                                        input->print_sensitivity_cmy.data, cmy);
        if (failed < 0) {
            std::memcpy(out->cmy, cmy, sizeof cmy);
            return status(FJ_STATUS_SUCCESS);
        }
        category = FJ_STATUS_PREPARATION_FAILURE;
        reason = FJ_PRINT_PREFLASH_REASON_CHANNEL;
        channel = static_cast<uint32_t>(failed);
    }
    if (diagnostic_valid) write_diagnostic(error, "preflash", "channel", reason, channel);
    if (valid_failure) *failure = FjPrintPreflashFailure{reason, channel};
    return status(category);
}

FJ_EXPORT FjStatus fjx_prepare_print_balance(const FjPrintBalanceInput* input, FjPrintBalance* out,
                                             FjPrintBalanceFailure* failure, FjErrorBuffer* error) {
    const bool valid_out = valid_extent(out, 1), valid_failure = valid_extent(failure, 1);
    if (valid_out) *out = FjPrintBalance{};
    if (valid_failure) *failure = FjPrintBalanceFailure{FJ_PRINT_BALANCE_REASON_NONE, kNoIndex};
    const bool diagnostic_valid = admit_diagnostic(error);
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT, reason = FJ_PRINT_BALANCE_REASON_RAW, channel = kNoIndex;
    if (valid_out && valid_failure && diagnostic_valid && valid_extent(input, 1) && input->normalization_mode <= 255 &&
        span_exact(input->film_base_density, 81) && span_exact(input->film_channel_density_cmy, 243) &&
        span_exact(input->print_sensitivity_rgb, 243) && span_exact(input->filtered_main_illuminant, 81)) {
        FjPrintBalance b{};
        category = FJ_STATUS_PREPARATION_FAILURE;
        int failed = balance_raw(*input, input->baseline_density_cmy, b.baseline_raw_rgb);
        if (failed >= 0) {
            reason = FJ_PRINT_BALANCE_REASON_BASELINE_RAW;
            channel = static_cast<uint32_t>(failed);
        } else if ((failed = balance_raw(*input, input->compensated_density_cmy, b.compensated_raw_rgb)) >= 0) {
            reason = FJ_PRINT_BALANCE_REASON_COMPENSATED_RAW;
            channel = static_cast<uint32_t>(failed);
        } else if (!balance_scalars(*input, b)) {
            reason = FJ_PRINT_BALANCE_REASON_BALANCE;
        } else {
            *out = b;
            return status(FJ_STATUS_SUCCESS);
        }
    }
    if (diagnostic_valid) write_diagnostic(error, "balance", "channel", reason, channel);
    if (valid_failure) *failure = FjPrintBalanceFailure{reason, channel};
    return status(category);
}

FJ_EXPORT FjStatus fjx_prepare_print_density(const FjPrintDensityInput* input, float* totals, float* layers,
                                             size_t capacity, uint64_t* print_hash, FjPrintDensityFailure* failure,
                                             FjErrorBuffer* error) {
    const bool valid_failure = valid_extent(failure, 1), valid_hash = valid_extent(print_hash, 1);
    if (valid_failure) *failure = FjPrintDensityFailure{FJ_PRINT_DENSITY_REASON_NONE, kNoIndex, kNoIndex, kNoIndex};
    if (valid_hash) *print_hash = 0;
    const bool diagnostic_valid = admit_diagnostic(error);
    const size_t count = input && valid_extent(input, 1) ? input->count : 0;
    const bool valid_totals = count && capacity >= count && count <= SIZE_MAX / 9 && valid_extent(totals, count * 3);
    const bool valid_layers = !layers || (valid_totals && valid_extent(layers, count * 9));
    if (valid_totals) std::memset(totals, 0, count * 3 * sizeof(float));
    if (valid_totals && layers && valid_layers) std::memset(layers, 0, count * 9 * sizeof(float));
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT, reason = FJ_PRINT_DENSITY_REASON_RAW;
    uint32_t index = kNoIndex, channel = kNoIndex, layer = kNoIndex;
    if (valid_failure && valid_hash && diagnostic_valid && valid_totals && valid_layers &&
        valid_extent(input->log_exposure, count) && input->polarity <= FJ_POLARITY_POSITIVE && input->reserved == 0) {
        category = FJ_STATUS_PREPARATION_FAILURE;
        const double gamma = input->gamma;
        if (!std::isfinite(gamma) || gamma <= 0.0) {
            reason = FJ_PRINT_DENSITY_REASON_INVALID_GAMMA;
        } else {
            const FjDensityCurveModel adjusted = print_gamma_model(input->model, gamma);
            const bool positive = input->polarity == FJ_POLARITY_POSITIVE;
            std::vector<float> t, l;
            try {
                t.resize(count * 3);
                l.resize(count * 9);
            } catch (const std::bad_alloc&) {
                if (valid_failure) *failure = FjPrintDensityFailure{reason, index, channel, layer};
                return status(FJ_STATUS_ALLOCATION_FAILURE);
            }
            DensityFault fault;
            size_t i = 0;
            for (; i < count; ++i) {
                fault = sample_density(adjusted, positive, input->log_exposure[i], &t[i * 3],
                                       reinterpret_cast<float(*)[3]>(&l[i * 9]));
                if (fault.reason != FJ_PRINT_DENSITY_REASON_NONE) break;
            }
            if (fault.reason != FJ_PRINT_DENSITY_REASON_NONE) {
                reason = fault.reason;
                index = static_cast<uint32_t>(i);
                channel = fault.channel;
                layer = fault.layer;
            } else {
                // PrintDensityCurves identity (sampled.rs:264-274): raw words, zero signs kept.
                uint64_t h = kFnvOffset;
                const uint64_t n = count;
                fnv_bytes(h, &n, 8);
                for (size_t k = 0; k < count; ++k) {
                    const float narrowed = static_cast<float>(input->log_exposure[k]);
                    fnv_bytes(h, &narrowed, 4);
                }
                fnv_bytes(h, t.data(), t.size() * sizeof(float));
                if (h == 0) {
                    reason = FJ_PRINT_DENSITY_REASON_ZERO_HASH;
                } else {
                    std::memcpy(totals, t.data(), t.size() * sizeof(float));
                    if (layers) std::memcpy(layers, l.data(), l.size() * sizeof(float));
                    *print_hash = h;
                    return status(FJ_STATUS_SUCCESS);
                }
            }
        }
    }
    if (diagnostic_valid && error && error->capacity) {
        char text[128];
        std::snprintf(text, sizeof text, "print density reason=%u index=%u channel=%u layer=%u", reason, index,
                      channel, layer);
        size_t n = std::strlen(text);
        if (n > error->capacity - 1) n = error->capacity - 1;
        std::memcpy(error->data, text, n);
        error->data[n] = '\0';
        error->length = n;
    }
    if (valid_failure) *failure = FjPrintDensityFailure{reason, index, channel, layer};
    return status(category);
}

}  // extern "C"
