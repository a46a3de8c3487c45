// fj_host_imports.cpp — the optional functions a host process defines for
// the render library (include/fj/juicer_diffusion_api.h,
// include/fj/juicer_dir_boundary_api.h), resolved at run time.
//
// The reference's native library links them from the Rust plug-in. This
// library looks them up with dlsym in the process's global scope:
//   * fj_diffusion_components_prepare, when defined, supplies the nine PSF
//     components of every diffusion stage (src/DiffusionHostBehavior.cpp:
//     52-82); otherwise the library expands the sample itself;
//   * fj_dir_boundary_describe, when defined, certifies the reflected
//     boundary of every recursive spatial-DIR component of a frame
//     (src/Cuda/JuicerCudaResources.cpp:6298-6316); a frame whose boundary the
//     host cannot certify is refused with the host's status and diagnostic.
// A host answer that breaks the call's contract (src/JuicerDirBoundary.cpp:
// 38-56, the components envelope of DiffusionHostBehavior.cpp) is refused as
// the reference refuses it.
#include <dlfcn.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_set>

#include "../../include/fj/juicer_diffusion_api.h"
#include "../../include/fj/juicer_dir_boundary_api.h"
#include "fj_diffusion_host.h"
#include "fj_host_imports.h"

namespace fjx {

namespace {

using ComponentsFn = FjStatus (*)(const FjDiffusionSample*, FjDiffusionComponents*, FjDiffusionFailure*, FjErrorBuffer*);
using DescribeFn = FjStatus (*)(const FjDirBoundaryInput*, uint64_t*, FjDirBoundaryFailure*, FjErrorBuffer*);

template <typename Fn>
Fn lookup(const char* name) {
    return reinterpret_cast<Fn>(dlsym(RTLD_DEFAULT, name));
}

bool cleared(const FjDiffusionComponents& c) {
    const auto* b = reinterpret_cast<const unsigned char*>(&c);
    for (size_t i = 0; i < sizeof(c); ++i)
        if (b[i]) return false;
    return true;
}

// The reference's texts of the boundary failure reasons (JuicerDirBoundary.cpp:19-33).
constexpr std::array<const char*, 14> kBoundaryMessages{
    "",
    "spatial DIR reflected-boundary arguments are invalid",
    "spatial DIR reflected-boundary request is invalid",
    "spatial DIR reflected-boundary axis size is invalid",
    "spatial DIR reflected-boundary policy is invalid",
    "spatial DIR causal reflected-boundary solve is singular",
    "spatial DIR causal reflected-boundary weights are nonfinite",
    "spatial DIR causal reflected-boundary certification failed",
    "spatial DIR reflected terminal system escaped its tail",
    "spatial DIR reflected terminal solve is singular",
    "spatial DIR reflected terminal certification failed",
    "spatial DIR reflected-boundary storage size is invalid",
    "spatial DIR reflected-boundary allocation failed",
    "spatial DIR reflected-boundary preparation panicked"};
// Status category each reason carries.
constexpr std::array<uint32_t, 14> kBoundaryCategories{0, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 4, 8};

}  // namespace

bool host_defines_diffusion_components() { return lookup<ComponentsFn>("fj_diffusion_components_prepare") != nullptr; }
bool host_defines_dir_boundary() { return lookup<DescribeFn>("fj_dir_boundary_describe") != nullptr; }

bool stage_components(const FjDiffusion& d, double lambda_px[kDiffusionComponents],
                      double weights[3][kDiffusionComponents], uint32_t& category, std::string& diagnostic) {
    const ComponentsFn host = lookup<ComponentsFn>("fj_diffusion_components_prepare");
    if (!host) {
        std::string field;
        if (diffusion_components(d, lambda_px, weights, field)) return true;
        category = FJ_STATUS_PREPARATION_FAILURE;
        diagnostic = "InvalidDiffusionHostBehavior field=" + field;
        return false;
    }
    FjDiffusionSample in{};
    in.family = d.family;
    for (int i = 0; i < 3; ++i) {
        in.group_weights[i] = d.group_weights[i];
        in.group_lambda_um[i] = d.group_lambda_um[i];
    }
    in.warmth = d.warmth;
    in.spatial_scale = d.spatial_scale;
    in.pixel_size_um = d.pixel_size_um;
    in.radius_px = d.radius_px;
    in.hash = d.sample_hash;
    FjDiffusionComponents out{};
    FjDiffusionFailure failure{};
    std::array<char, 256> text{};
    FjErrorBuffer error{text.data(), text.size(), 0};
    const FjStatus st = host(&in, &out, &failure, &error);
    const bool text_ok = error.length < text.size() && text[error.length] == '\0';
    const bool success = st.category == FJ_STATUS_SUCCESS;
    if (!text_ok || st.api != FJ_API_NONE || st.native_code != 0 || success != (failure.reason == 0) ||
        failure.reason > FJ_DIFFUSION_REASON_PANIC || (!success && !cleared(out)) || (success && error.length != 0)) {
        category = FJ_STATUS_PREPARATION_FAILURE;
        diagnostic = "InvalidDiffusionHostBehavior field=components";
        return false;
    }
    if (!success) {
        category = st.category;
        diagnostic.assign(text.data(), error.length);
        return false;
    }
    for (int k = 0; k < kDiffusionComponents; ++k) {
        lambda_px[k] = out.lambda_px[k];
        weights[0][k] = out.red_weights[k];
        weights[1][k] = out.green_weights[k];
        weights[2][k] = out.blue_weights[k];
    }
    // The same admission of the expansion the native path applies
    // (valid_components, JuicerCudaDiffusion.cu:102-126).
    double sums[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < kDiffusionComponents; ++k) {
        bool ok = std::isfinite(lambda_px[k]) && lambda_px[k] > 0.0;
        for (int c = 0; c < 3; ++c) {
            ok = ok && std::isfinite(weights[c][k]) && weights[c][k] >= 0.0;
            sums[c] += weights[c][k];
        }
        if (!ok) {
            category = FJ_STATUS_PREPARATION_FAILURE;
            diagnostic = "InvalidDiffusionHostBehavior field=component_output";
            return false;
        }
    }
    for (double s : sums)
        if (!(std::isfinite(s) && s > 0.0)) {
            category = FJ_STATUS_PREPARATION_FAILURE;
            diagnostic = "InvalidDiffusionHostBehavior field=component_output";
            return false;
        }
    return true;
}

namespace {

using AcquireFn = FjStatus (*)(const FjDirBoundaryInput*, FjDirBoundaryOwner**, FjDirBoundaryFailure*, FjErrorBuffer*);
using ViewFn = FjStatus (*)(const FjDirBoundaryOwner*, FjDirBoundaryView*, FjDirBoundaryFailure*, FjErrorBuffer*);
using ReleaseFn = FjStatus (*)(FjDirBoundaryOwner*, FjErrorBuffer*);

// Identities of boundaries the host certified in this process.
std::mutex g_certified_mu;
std::unordered_set<uint64_t> g_certified;

// The reference's envelope of a boundary call (JuicerDirBoundary.cpp:38-56).
bool boundary_envelope(const FjDirBoundaryInput& in, const FjStatus& st, const FjDirBoundaryFailure& f,
                       const FjErrorBuffer& error, const std::array<char, 256>& text) {
    if (f.reason >= kBoundaryCategories.size() || st.category != kBoundaryCategories[f.reason] ||
        st.api != FJ_API_NONE || st.native_code != 0 || error.length >= text.size() || text[error.length] != '\0' ||
        std::strcmp(text.data(), kBoundaryMessages[f.reason]) != 0)
        return false;
    if (f.axis > 2 || (f.axis == 0 ? f.length != 0 : f.length != (f.axis == 1 ? in.width : in.height))) return false;
    if (f.reason == FJ_DIR_BOUNDARY_CAUSAL_CERTIFICATION)
        return f.axis != 0 && f.row < 3 && std::memcmp(&f.sigma_px, &in.component.sigma_px, sizeof(float)) == 0 &&
               std::memcmp(&f.tolerance, &in.component.boundary_certification_tolerance, sizeof(double)) == 0;
    const uint32_t zero32 = 0;
    const uint64_t zero64 = 0;
    if (f.row != 3 || std::memcmp(&f.sigma_px, &zero32, 4) != 0 || std::memcmp(&f.sum, &zero64, 8) != 0 ||
        std::memcmp(&f.tolerance, &zero64, 8) != 0)
        return false;
    if (f.reason <= 2 || f.reason == 11 || f.reason == 13) return f.axis == 0;
    return f.reason == 12 || f.axis != 0;
}

std::string boundary_text(const FjDirBoundaryFailure& f) {
    if (f.reason != FJ_DIR_BOUNDARY_CAUSAL_CERTIFICATION) return kBoundaryMessages[f.reason];
    std::ostringstream detail;
    detail << kBoundaryMessages[f.reason] << " length=" << f.length << " sigma=" << f.sigma_px << " sum=" << f.sum
           << " tolerance=" << f.tolerance;
    return detail.str();
}

}  // namespace

bool host_dir_boundaries(const FjSpatialDir& sd, int width, int height, uint32_t& category, std::string& diagnostic) {
    const DescribeFn describe = lookup<DescribeFn>("fj_dir_boundary_describe");
    if (!describe) return true;
    const AcquireFn acquire = lookup<AcquireFn>("fj_dir_boundary_acquire");
    const ViewFn view = lookup<ViewFn>("fj_dir_boundary_view");
    const ReleaseFn release = lookup<ReleaseFn>("fj_dir_boundary_release");
    auto invalid = [&]() {
        category = FJ_STATUS_PREPARATION_FAILURE;
        diagnostic = "spatial DIR reflected-boundary view is invalid";
        return false;
    };
    auto refuse = [&](const FjStatus& st, const FjDirBoundaryFailure& f) {
        category = st.category;
        diagnostic = boundary_text(f);
        return false;
    };
    if (!acquire || !view || !release) return invalid();
    for (size_t i = 0; i < sd.component_count && i < 4; ++i) {
        const FjDirGaussian& g = sd.components[i];
        if (static_cast<uint8_t>(g.reference_operator) != FJ_DIR_OPERATOR_YVV_REFLECT || !(g.weight > 0.0f)) continue;
        FjDirBoundaryInput in{};
        in.component = g;
        in.width = width;
        in.height = height;
        uint64_t key = 0;
        FjDirBoundaryFailure f{};
        std::array<char, 256> text{};
        FjErrorBuffer error{text.data(), text.size(), 0};
        FjStatus st = describe(&in, &key, &f, &error);
        if (!boundary_envelope(in, st, f, error, text) || (f.reason != 0 && f.reason != 1 && f.reason != 2 && f.reason != 13) ||
            (st.category != FJ_STATUS_SUCCESS && key != 0))
            return invalid();
        if (st.category != FJ_STATUS_SUCCESS) return refuse(st, f);
        {
            std::lock_guard<std::mutex> lock(g_certified_mu);
            if (g_certified.count(key)) continue;
        }
        // Cold identity: the host solves and certifies the boundary.
        FjDirBoundaryOwner* owner = nullptr;
        f = FjDirBoundaryFailure{};
        text.fill('\0');
        error = FjErrorBuffer{text.data(), text.size(), 0};
        st = acquire(&in, &owner, &f, &error);
        if (!boundary_envelope(in, st, f, error, text) || (st.category == FJ_STATUS_SUCCESS ? !owner : owner != nullptr)) {
            if (owner) (void)release(owner, nullptr);
            return invalid();
        }
        if (st.category != FJ_STATUS_SUCCESS) return refuse(st, f);
        FjDirBoundaryView v{};
        f = FjDirBoundaryFailure{};
        text.fill('\0');
        error = FjErrorBuffer{text.data(), text.size(), 0};
        st = view(owner, &v, &f, &error);
        const bool ok = boundary_envelope(in, st, f, error, text) && (f.reason == 0 || f.reason == 1 || f.reason == 13);
        const bool admitted = ok && st.category == FJ_STATUS_SUCCESS && v.identity == key && v.width == width &&
                              v.height == height && v.weights != nullptr &&
                              v.count == 3 * (static_cast<size_t>(v.horizontal_length) + v.vertical_length);
        (void)release(owner, nullptr);
        if (!ok || (st.category == FJ_STATUS_SUCCESS && !admitted)) return invalid();
        if (st.category != FJ_STATUS_SUCCESS) return refuse(st, f);
        std::lock_guard<std::mutex> lock(g_certified_mu);
        g_certified.insert(key);
    }
    return true;
}

}  // namespace fjx
