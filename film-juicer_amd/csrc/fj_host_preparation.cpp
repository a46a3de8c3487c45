// fj_host_preparation.cpp — the library's own builds of the two products a
// host supplies through FjHostPreparation (native/juicer_cuda_call_scope.h:
// 12-16): the PCHIP scanner LUT (rust scanner/lut.rs:186-299) and normalized
// Gaussian FIR weights (rust optics/gaussian.rs:18-56), exported with the
// callbacks' signatures and failure records (reasons and categories of
// rust/film-juicer-plugin/src/optics_bridge.rs and the validation in
// native/juicer_host_preparation.cpp:28-109).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/fj/fj_mi355x_ext.h"
#include "fj_host.h"
#include "fj_owner.h"
#include "fj_spatial.h"

#define FJ_EXPORT __attribute__((visibility("default")))

using namespace fjx;

namespace {

// Storage of the last successful build on each thread: the views borrow it
// until the thread's next call.
thread_local ScannerLut t_lut;
thread_local std::vector<float> t_weights;

FjFloatSpan view_of(const std::vector<float>& v) { return FjFloatSpan{v.data(), v.size()}; }

bool span_is(const FjFloatSpan& s, size_t count) { return s.count == count && (count == 0 || s.data != nullptr); }

}  // namespace

extern "C" {

FJ_EXPORT FjStatus fjx_prepare_scanner_lut(void* /*user*/, const FjScannerLutInput* in, FjScannerLutView* view,
                                           FjScannerLutFailure* failure, FjErrorBuffer* error) {
    if (view) *view = FjScannerLutView{};
    if (failure) *failure = FjScannerLutFailure{0u, UINT32_MAX, UINT64_MAX, 0u};
    auto reject = [&](uint32_t reason, uint32_t category, const std::string& text, uint32_t channel = UINT32_MAX,
                      uint64_t index = UINT64_MAX) {
        if (failure) *failure = FjScannerLutFailure{reason, channel, index, 0u};
        return fail(error, category, text);
    };
    try {
        if (!in || !view || !failure || bad_error_buffer(error))
            return reject(FJ_SCANNER_LUT_REASON_RAW, FJ_STATUS_UNSUPPORTED_INPUT, "invalid scanner LUT arguments");
        if (in->resolution < 2 || in->resolution > 256)
            return reject(FJ_SCANNER_LUT_REASON_RESOLUTION, FJ_STATUS_PREPARATION_FAILURE,
                          "canonical scan LUT resolution must be at least two (and at most 256 here)");
        const int k = in->sample_count;
        if (k > 0) {
            const size_t n = static_cast<size_t>(k);
            for (const FjFloatSpan* sp : {&in->eps_c, &in->eps_m, &in->eps_y, &in->weighted_x, &in->weighted_y,
                                          &in->weighted_z})
                if (!span_is(*sp, n))
                    return reject(FJ_SCANNER_LUT_REASON_RAW, FJ_STATUS_UNSUPPORTED_INPUT, "invalid scanner LUT arguments");
            if (in->has_baseline && !span_is(in->baseline_min, n))
                return reject(FJ_SCANNER_LUT_REASON_RAW, FJ_STATUS_UNSUPPORTED_INPUT, "invalid scanner LUT arguments");
        }
        ScannerLutInput li{};
        li.eps_c = in->eps_c.data;
        li.eps_m = in->eps_m.data;
        li.eps_y = in->eps_y.data;
        li.weighted_x = in->weighted_x.data;
        li.weighted_y = in->weighted_y.data;
        li.weighted_z = in->weighted_z.data;
        li.baseline = in->has_baseline ? in->baseline_min.data : nullptr;
        li.sample_count = k;
        li.inv_yn = in->inv_yn;
        for (int c = 0; c < 3; ++c) {
            li.range_min_cmy[c] = in->range_min_cmy[c];
            li.inverse_span_cmy[c] = in->inverse_span_cmy[c];
        }
        li.medium = static_cast<int>(in->medium);
        li.resolution = static_cast<int>(in->resolution);
        std::string diag;
        uint64_t node = UINT64_MAX;
        uint32_t channel = UINT32_MAX;
        ScannerLut lut;
        if (!build_scanner_lut(li, lut, diag, &node, &channel))
            return reject(FJ_SCANNER_LUT_REASON_NONFINITE, FJ_STATUS_PREPARATION_FAILURE, diag, channel, node);
        t_lut = std::move(lut);
        view->resolution = static_cast<uint32_t>(t_lut.resolution);
        view->log2_xyz = view_of(t_lut.log2_xyz);
        view->slope_c = view_of(t_lut.slope_c);
        view->slope_m = view_of(t_lut.slope_m);
        view->slope_y = view_of(t_lut.slope_y);
        view->cell_min = view_of(t_lut.cell_min);
        view->cell_max = view_of(t_lut.cell_max);
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (const std::bad_alloc&) {
        if (view) *view = FjScannerLutView{};
        return reject(FJ_SCANNER_LUT_REASON_ALLOCATION, FJ_STATUS_ALLOCATION_FAILURE, "scanner LUT allocation failed");
    } catch (...) {
        if (view) *view = FjScannerLutView{};
        return reject(FJ_SCANNER_LUT_REASON_PANIC, FJ_STATUS_INTERNAL_FAILURE, "scanner LUT preparation failed");
    }
}

FJ_EXPORT FjStatus fjx_prepare_gaussian(void* /*user*/, const FjGaussianKernelInput* in, FjGaussianKernelView* view,
                                        FjGaussianKernelFailure* failure, FjErrorBuffer* error) {
    if (view) *view = FjGaussianKernelView{};
    if (failure) failure->reason = 0;
    auto reject = [&](uint32_t reason, uint32_t category, const char* text) {
        if (failure) failure->reason = reason;
        return fail(error, category, text);
    };
    try {
        if (!in || !view || !failure || bad_error_buffer(error))
            return reject(1, FJ_STATUS_UNSUPPORTED_INPUT, "invalid Gaussian kernel arguments");
        if (in->radius < 1 || in->radius > 2048)
            return reject(2, FJ_STATUS_PREPARATION_FAILURE, "Gaussian kernel radius outside supported range");
        if (!std::isfinite(in->sigma_px) || !(in->sigma_px > 0.0f))
            return reject(3, FJ_STATUS_PREPARATION_FAILURE, "invalid Gaussian kernel sigma");
        std::vector<float> w = gaussian_weights(in->radius, in->sigma_px);
        if (w.size() != static_cast<size_t>(in->radius) * 2 + 1)
            return reject(3, FJ_STATUS_PREPARATION_FAILURE, "invalid Gaussian kernel sigma");
        t_weights = std::move(w);
        view->radius = in->radius;
        view->sigma_px = in->sigma_px;
        view->weights = view_of(t_weights);
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (const std::bad_alloc&) {
        if (view) *view = FjGaussianKernelView{};
        return reject(4, FJ_STATUS_ALLOCATION_FAILURE, "Gaussian kernel allocation failed");
    } catch (...) {
        if (view) *view = FjGaussianKernelView{};
        return reject(5, FJ_STATUS_INTERNAL_FAILURE, "Gaussian kernel preparation panicked");
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The frame's host preparation: when a call scope's render supplies
// FjHostPreparation callbacks, the frame's scanner LUT and Gaussian kernels
// come from them (native/juicer_host_preparation.cpp:28-109); without them
// the builds above are used.
// ---------------------------------------------------------------------------
namespace fjx {

namespace {

thread_local const FjHostPreparation* t_host = nullptr;
thread_local bool t_failed = false;
thread_local uint32_t t_category = 0;
thread_local std::string t_text;

void set_failure(uint32_t category, const std::string& text) {
    if (t_failed) return;  // the first failure of the frame is reported
    t_failed = true;
    t_category = category;
    t_text = text;
}

bool text_valid(const FjErrorBuffer& error, const char* text, size_t capacity) {
    return error.length < capacity && text[error.length] == '\0';
}

bool span_valid(const FjFloatSpan& s, size_t count) {
    return count && s.count == count && s.data && reinterpret_cast<uintptr_t>(s.data) % alignof(float) == 0;
}

}  // namespace

HostPreparationScope::HostPreparationScope(const FjHostPreparation* host) : previous_(t_host) {
    t_host = host;
    t_failed = false;
}

HostPreparationScope::~HostPreparationScope() { t_host = previous_; }

bool take_host_failure(uint32_t& category, std::string& text) {
    if (!t_failed) return false;
    t_failed = false;
    category = t_category;
    text = t_text;
    return true;
}

std::vector<float> kernel_weights(int radius, float sigma_px) {
    if (!t_host) return gaussian_weights(radius, sigma_px);
    const FjHostPreparation& hp = *t_host;
    auto invalid = [&]() {
        set_failure(FJ_STATUS_PREPARATION_FAILURE, "Gaussian kernel projection invalid");
        return std::vector<float>{};
    };
    if (!hp.user || !hp.scanner_lut || !hp.gaussian) return invalid();
    FjGaussianKernelInput in{radius, sigma_px};
    FjGaussianKernelView view{};
    FjGaussianKernelFailure failure{};
    char text[256] = {};
    FjErrorBuffer error{text, sizeof(text), 0};
    const FjStatus st = hp.gaussian(hp.user, &in, &view, &failure, &error);
    static const uint32_t kCategories[6] = {FJ_STATUS_SUCCESS,             FJ_STATUS_UNSUPPORTED_INPUT,
                                            FJ_STATUS_PREPARATION_FAILURE, FJ_STATUS_PREPARATION_FAILURE,
                                            FJ_STATUS_ALLOCATION_FAILURE,  FJ_STATUS_INTERNAL_FAILURE};
    if (!text_valid(error, text, sizeof(text)) || st.api != FJ_API_NONE || st.native_code != 0 || failure.reason >= 6 ||
        st.category != kCategories[failure.reason])
        return invalid();
    if (st.category != FJ_STATUS_SUCCESS) {
        uint32_t sigma_bits = 0;
        std::memcpy(&sigma_bits, &view.sigma_px, 4);
        if (view.radius || sigma_bits || view.weights.data || view.weights.count) return invalid();
        set_failure(st.category, std::string(text, error.length));
        return {};
    }
    const size_t count = radius >= 1 && radius <= 2048 ? static_cast<size_t>(radius) * 2 + 1 : 0;
    if (error.length || view.radius != radius || std::memcmp(&view.sigma_px, &sigma_px, 4) != 0 ||
        !span_valid(view.weights, count))
        return invalid();
    return std::vector<float>(view.weights.data, view.weights.data + count);
}

bool frame_scanner_lut(const ScannerLutInput& in, ScannerLut& out, std::string& diagnostic) {
    if (!t_host) return build_scanner_lut(in, out, diagnostic);
    const FjHostPreparation& hp = *t_host;
    auto invalid = [&]() {
        set_failure(FJ_STATUS_INTERNAL_FAILURE, "Scanner LUT preparation: invalid transport contract");
        diagnostic = "invalid transport contract";
        return false;
    };
    if (!hp.user || !hp.scanner_lut || !hp.gaussian) return invalid();
    const size_t k = in.sample_count > 0 ? static_cast<size_t>(in.sample_count) : 0;
    auto sp = [k](const float* p) { return FjFloatSpan{p, p ? k : 0}; };
    FjScannerLutInput ci{};
    ci.eps_c = sp(in.eps_c);
    ci.eps_m = sp(in.eps_m);
    ci.eps_y = sp(in.eps_y);
    ci.weighted_x = sp(in.weighted_x);
    ci.weighted_y = sp(in.weighted_y);
    ci.weighted_z = sp(in.weighted_z);
    ci.baseline_min = sp(in.baseline);
    for (int c = 0; c < 3; ++c) {
        ci.range_min_cmy[c] = in.range_min_cmy[c];
        ci.inverse_span_cmy[c] = in.inverse_span_cmy[c];
    }
    ci.sample_count = in.sample_count;
    ci.inv_yn = in.inv_yn;
    ci.has_baseline = in.baseline ? 1u : 0u;
    ci.medium = static_cast<uint32_t>(in.medium);
    ci.resolution = static_cast<uint32_t>(in.resolution);
    FjScannerLutView view{};
    FjScannerLutFailure failure{};
    char text[256] = {};
    FjErrorBuffer error{text, sizeof(text), 0};
    const FjStatus st = hp.scanner_lut(hp.user, &ci, &view, &failure, &error);
    if (!text_valid(error, text, sizeof(text)) || st.api != FJ_API_NONE || st.native_code != 0) return invalid();
    const FjFloatSpan arrays[6] = {view.log2_xyz, view.slope_c, view.slope_m, view.slope_y, view.cell_min, view.cell_max};
    if (st.category != FJ_STATUS_SUCCESS) {
        const uint32_t r = failure.reason;
        const uint32_t want = r == FJ_SCANNER_LUT_REASON_RAW                                            ? FJ_STATUS_UNSUPPORTED_INPUT
                              : r == FJ_SCANNER_LUT_REASON_SIZE || r == FJ_SCANNER_LUT_REASON_ALLOCATION ? FJ_STATUS_ALLOCATION_FAILURE
                              : r == FJ_SCANNER_LUT_REASON_PANIC                                         ? FJ_STATUS_INTERNAL_FAILURE
                                                                                                         : FJ_STATUS_PREPARATION_FAILURE;
        bool empty = view.resolution == 0;
        for (const FjFloatSpan& a : arrays) empty = empty && !a.data && !a.count;
        const bool located = r == FJ_SCANNER_LUT_REASON_NONFINITE ? failure.channel <= 2 && failure.index != UINT64_MAX
                                                                   : failure.channel == UINT32_MAX && failure.index == UINT64_MAX;
        if (r < 1 || r > 6 || st.category != want || !empty || !located) return invalid();
        set_failure(st.category, std::string(text, error.length));
        diagnostic = std::string(text, error.length);
        return false;
    }
    if (error.length || failure.reason || failure.channel != UINT32_MAX || failure.index != UINT64_MAX || failure.buffer ||
        view.resolution != ci.resolution || ci.resolution < 2)
        return invalid();
    const size_t r = ci.resolution, nodes = r * r * r * 3, cells = (r - 1) * (r - 1) * (r - 1) * 3;
    for (int i = 0; i < 6; ++i)
        if (!span_valid(arrays[i], i < 4 ? nodes : cells)) return invalid();
    out.resolution = static_cast<int>(r);
    std::vector<float>* dst[6] = {&out.log2_xyz, &out.slope_c, &out.slope_m, &out.slope_y, &out.cell_min, &out.cell_max};
    for (int i = 0; i < 6; ++i) dst[i]->assign(arrays[i].data, arrays[i].data + arrays[i].count);
    return true;
}

}  // namespace fjx
