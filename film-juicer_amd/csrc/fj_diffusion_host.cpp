// fj_diffusion_host.cpp — camera/enlarger diffusion: descriptor rules, PSF
// components, transfer tables and the per-frame stage.
//
// Descriptor rules follow decode_optics / decode_diffusion
// (native/juicer_cuda_descriptors.cpp:183-243); the components follow the
// host behaviour the native side calls (admit_components, expand_group,
// halo_weights and Components::new, rust optics/diffusion_frame.rs:161-420);
// the PSF samples follow build_wrapped_psf_kernel / normalize
// (src/Cuda/Diffusion/JuicerCudaDiffusion.cu:276-400): an FP64 sum of nine
// exponentials rounded to float, divided by the FP64 sum of the rounded
// samples.
#include "fj_diffusion_host.h"
#include "fj_host_imports.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fj_diffusion.h"
#include "fj_spatial.h"

namespace fjx {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr uint64_t kFnvOffset = 0xCBF29CE484222325ull;
constexpr uint64_t kFnvPrime = 0x100000001B3ull;

void fnv(uint64_t& h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= kFnvPrime;
    }
}

void fnv_double(uint64_t& h, double v) {
    uint64_t bits = 0;
    if (v != 0.0) std::memcpy(&bits, &v, 8);  // both zeros hash as +0
    unsigned char b[8];
    for (int i = 0; i < 8; ++i) b[i] = static_cast<unsigned char>(bits >> (8 * i));
    fnv(h, b, 8);
}

// Sample identity (sample_hash, diffusion_frame.rs:161-180).
uint64_t sample_hash(const FjDiffusion& d) {
    uint64_t h = kFnvOffset;
    static const char tag[] = "diffusion-psf-sample";
    fnv(h, tag, sizeof(tag));  // includes the terminating zero
    const unsigned char version[4] = {1, 0, 0, 0};
    fnv(h, version, 4);
    const unsigned char family = static_cast<unsigned char>(d.family);
    fnv(h, &family, 1);
    for (double v : d.group_weights) fnv_double(h, v);
    for (double v : d.group_lambda_um) fnv_double(h, v);
    fnv_double(h, d.warmth);
    fnv_double(h, d.spatial_scale);
    fnv_double(h, d.pixel_size_um);
    const uint32_t r = static_cast<uint32_t>(d.radius_px);
    const unsigned char rb[4] = {static_cast<unsigned char>(r), static_cast<unsigned char>(r >> 8),
                                 static_cast<unsigned char>(r >> 16), static_cast<unsigned char>(r >> 24)};
    fnv(h, rb, 4);
    return h;
}

double ordered_max(double v, double floor) { return v < floor ? floor : v; }

struct Group {
    double lambda_um[4];
    double weights[4];
    int count;
};

bool expand_group(double center, double spread, int count, const double* alpha, Group& g) {
    g.count = count;
    const double lo = std::log(center / spread), hi = std::log(center * spread);
    if (!std::isfinite(lo) || !std::isfinite(hi)) return false;
    for (int i = 0; i < count; ++i) {
        double ll = lo;
        if (i + 1 == count)
            ll = hi;
        else if (i != 0)
            ll = lo + (hi - lo) * (static_cast<double>(i) / (count - 1));
        g.lambda_um[i] = std::exp(ll);
        g.weights[i] = alpha ? std::pow(g.lambda_um[i], 2.0 - *alpha) : 1.0;
        if (!(std::isfinite(g.lambda_um[i]) && g.lambda_um[i] > 0.0 && std::isfinite(g.weights[i]) && g.weights[i] >= 0.0))
            return false;
    }
    double total = 0.0;
    for (int i = 0; i < count; ++i) total += g.weights[i];
    if (!(std::isfinite(total) && total > 0.0)) return false;
    for (int i = 0; i < count; ++i) g.weights[i] /= total;
    return true;
}

}  // namespace

bool diffusion_components(const FjDiffusion& d, double lambda_px[kDiffusionComponents],
                          double weights[3][kDiffusionComponents], std::string& field) {
    // admit_components (diffusion_frame.rs:223-262)
    if (d.sample_hash == 0) return field = "sample_descriptor_inactive", false;
    if (d.family > 3) return field = "family", false;
    double total = 0.0;
    for (double w : d.group_weights) {
        if (!std::isfinite(w) || w < 0.0) return field = "group_weights", false;
        total += w;
    }
    if (!(std::isfinite(total) && total > 0.0)) return field = "group_weights", false;
    for (double c : d.group_lambda_um)
        if (!(std::isfinite(c) && c > 0.0)) return field = "group_centers_um", false;
    if (!std::isfinite(d.warmth) || d.warmth < -1.5 || d.warmth > 1.5) return field = "effective_warmth", false;
    if (!(std::isfinite(d.spatial_scale) && d.spatial_scale > 0.0)) return field = "spatial_scale", false;
    if (!(std::isfinite(d.pixel_size_um) && d.pixel_size_um > 0.0)) return field = "pixel_size_um", false;
    if (d.radius_px <= 0) return field = "radius_pixels", false;
    if (sample_hash(d) != d.sample_hash) return field = "sample_descriptor_hash", false;

    static const double kBloomAlpha[4] = {3.2, 3.5, 2.9, 2.5};
    Group groups[3];
    if (!expand_group(d.group_lambda_um[0], 1.5, 2, nullptr, groups[0]) ||
        !expand_group(d.group_lambda_um[1], 2.0, 3, nullptr, groups[1]) ||
        !expand_group(d.group_lambda_um[2], 2.5, 4, &kBloomAlpha[d.family], groups[2]))
        return field = "component_expansion", false;

    // Halo warmth: per-channel reweighting along a centred gradient (halo_weights, :335-373).
    const Group& halo = groups[1];
    double grad[3], target = 0.0, wg = 0.0;
    for (int i = 0; i < 3; ++i) {
        grad[i] = -1.0 + 2.0 * i / 2.0;
        target += halo.weights[i];
        wg += halo.weights[i] * grad[i];
    }
    if (!(std::isfinite(target) && target > 0.0)) return field = "halo_channel_weights", false;
    wg /= target;
    for (double& g : grad) g -= wg;
    static const double kAxis[3] = {1.30, 0.15, -1.45};
    double halo_rgb[3][3];
    for (int ch = 0; ch < 3; ++ch) {
        double t = 0.0;
        for (int i = 0; i < 3; ++i) {
            halo_rgb[ch][i] = ordered_max(halo.weights[i] * (1.0 + d.warmth * kAxis[ch] * grad[i]), 0.0);
            t += halo_rgb[ch][i];
        }
        if (t > 0.0) {
            const double s = target / t;
            for (double& v : halo_rgb[ch]) v *= s;
        } else {
            for (int i = 0; i < 3; ++i) halo_rgb[ch][i] = halo.weights[i];
        }
    }
    const double scale = ordered_max(d.spatial_scale, 1e-6);
    int k = 0;
    for (int g = 0; g < 3; ++g)
        for (int i = 0; i < groups[g].count; ++i, ++k) {
            lambda_px[k] = ordered_max(groups[g].lambda_um[i] * scale / d.pixel_size_um, 1e-6);
            for (int ch = 0; ch < 3; ++ch)
                weights[ch][k] = g == 1 ? d.group_weights[1] * halo_rgb[ch][i] : d.group_weights[g] * groups[g].weights[i];
        }
    double sums[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < kDiffusionComponents; ++j) {
        if (!(std::isfinite(lambda_px[j]) && lambda_px[j] > 0.0)) return field = "component_output", false;
        for (int ch = 0; ch < 3; ++ch) {
            if (!std::isfinite(weights[ch][j]) || weights[ch][j] < 0.0) return field = "component_output", false;
            sums[ch] += weights[ch][j];
        }
    }
    // valid_components (JuicerCudaDiffusion.cu:102-126)
    for (double s : sums)
        if (!(std::isfinite(s) && s > 0.0)) return field = "component_output", false;
    return true;
}

bool diffusion_decode(const FjOptics& o, bool print_route, std::string& field) {
    const bool camera = (o.flags & FJ_OPTICS_CAMERA_DIFFUSION) != 0;
    const bool enlarger = (o.flags & FJ_OPTICS_ENLARGER_DIFFUSION) != 0;
    if (enlarger && !print_route) return field = "optics.flags", false;
    if (!camera && !enlarger) {
        if (o.diffusion_hash != 0) return field = "optics.inactive_diffusion", false;
        return true;
    }
    if (o.diffusion_hash == 0) return field = "optics.diffusion_hash", false;
    for (int s = 0; s < 2; ++s) {
        if (!(s == 0 ? camera : enlarger)) continue;
        const FjDiffusion& d = s == 0 ? o.camera : o.enlarger;
        if (d.family > 3 || d.hash == 0 || d.sample_hash == 0) return field = "optics.diffusion", false;
    }
    return true;
}

bool diffusion_quadrant(const FjDiffusion& d, int channel, std::vector<double>& q) {
    double lam[kDiffusionComponents], w[3][kDiffusionComponents];
    uint32_t category = 0;
    std::string diag;
    if (channel < 0 || channel > 2 || !stage_components(d, lam, w, category, diag)) return false;
    const int r = d.radius_px, n = r + 1;
    const double two_pi = 2.0 * kPi;
    q.assign(static_cast<size_t>(n) * n, 0.0);
    // The samples depend on dx^2 + dy^2 only: evaluate one octant, mirror.
    double sum = 0.0;
    for (int dy = 0; dy <= r; ++dy)
        for (int dx = dy; dx <= r; ++dx) {
            const double rad = std::sqrt(static_cast<double>(dx) * dx + static_cast<double>(dy) * dy);
            double v = 0.0;
            for (int k = 0; k < kDiffusionComponents; ++k)
                v += w[channel][k] * std::exp(-rad / lam[k]) / (two_pi * lam[k] * lam[k]);
            const double stored = static_cast<double>(static_cast<float>(v));
            q[static_cast<size_t>(dy) * n + dx] = stored;
            q[static_cast<size_t>(dx) * n + dy] = stored;
            const double copies = (dx == 0 ? 1.0 : 2.0) * (dy == 0 ? 1.0 : 2.0) * (dx == dy ? 1.0 : 2.0);
            sum += copies * stored;
        }
    const bool ok = std::isfinite(sum) && sum > 0.0;
    for (double& v : q) v = ok ? static_cast<double>(static_cast<float>(v / sum)) : std::nan("");
    return ok;
}

void diffusion_padded(const FjDiffusion& d, int width, int height, int& pw, int& ph) {
    pw = inplace_fft_length(width + 2 * d.radius_px);
    ph = inplace_fft_length(height + 2 * d.radius_px);
}

size_t diffusion_build_floats(const FjDiffusion& d, int pw, int ph) {
    const size_t k = static_cast<size_t>(d.radius_px) + 1;
    // C_W (k x pw), C_H (k x ph), M (3 x k x pw), all FP64, plus alignment.
    return 2 * (k * pw + k * ph + 3 * k * pw) + 64;
}

hipError_t diffusion_prepare(DeviceState& dev, const FjDiffusion& d, int width, int height, float* scratch,
                             hipStream_t stream, DiffusionStage& out) {
    out = DiffusionStage{};
    out.radius = d.radius_px;
    out.fraction = static_cast<float>(d.scatter_fraction);
    diffusion_padded(d, width, height, out.pw, out.ph);
    const int r = d.radius_px, k = r + 1, pw = out.pw, ph = out.ph;
    const uint64_t ident[3] = {d.sample_hash, static_cast<uint64_t>(pw), static_cast<uint64_t>(ph)};
    const uint64_t key = hash_bytes(ident, sizeof(ident), 0xD1FFull);
    auto cached = dev.transfers.find(key);
    if (cached != dev.transfers.end()) {
        // Keep the tables pinned for this frame.
        const void* ptr = nullptr;
        for (int c = 0; c < 3; ++c) {
            hipError_t e = dev.table(key ^ (0x100ull * (c + 1)), nullptr, 0, &ptr);
            if (e != hipSuccess) return e;
        }
        out.transfer = cached->second;
        return hipSuccess;
    }
    std::vector<double> q[3];
    for (int c = 0; c < 3; ++c)
        if (!diffusion_quadrant(d, c, q[c])) return hipErrorInvalidValue;
    std::vector<double> qs(static_cast<size_t>(3) * k * k);
    for (int c = 0; c < 3; ++c) std::copy(q[c].begin(), q[c].end(), qs.begin() + static_cast<size_t>(c) * k * k);
    const void* ptr = nullptr;
    hipError_t e = dev.table(key ^ 0x51ull, qs.data(), qs.size() * sizeof(double), &ptr);
    if (e != hipSuccess) return e;
    const double* qd = static_cast<const double*>(ptr);
    const double* cos_tab[2];
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis == 0 ? pw : ph;
        const uint64_t ck = 0xC05D000000000000ull ^ static_cast<uint64_t>(n);
        if (dev.tables.count(ck) == 0) {
            std::vector<double> t(2 * static_cast<size_t>(n));
            for (size_t m = 0; m < t.size(); ++m) t[m] = std::cos(kPi * static_cast<double>(m) / n);
            e = dev.table(ck, t.data(), t.size() * sizeof(double), &ptr);
        } else {
            e = dev.table(ck, nullptr, 0, &ptr);
        }
        if (e != hipSuccess) return e;
        cos_tab[axis] = static_cast<const double*>(ptr);
    }
    // Scratch: C_W, C_H, M (FP64, 8-byte aligned).
    double* base = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(scratch) + 255) & ~uintptr_t(255));
    double* cw = base;
    double* chh = cw + static_cast<size_t>(k) * pw;
    double* m = chh + static_cast<size_t>(k) * ph;
    DiffCosArgs ca{cos_tab[0], cw, k, pw};
    e = launch_diffusion_cos(ca, stream);
    if (e != hipSuccess) return e;
    ca = DiffCosArgs{cos_tab[1], chh, k, ph};
    e = launch_diffusion_cos(ca, stream);
    if (e != hipSuccess) return e;
    AtbArgs<double> s1{};
    for (int c = 0; c < 3; ++c) {
        s1.L[c] = qd + static_cast<size_t>(c) * k * k;
        s1.out[c] = m + static_cast<size_t>(c) * k * pw;
    }
    s1.R = cw;
    s1.M = k;
    s1.N = pw;
    s1.K = k;
    e = launch_diffusion_atb(s1, 3, stream);
    if (e != hipSuccess) return e;
    AtbArgs<float> s2{};
    DctTransfer tf{};
    for (int c = 0; c < 3; ++c) {
        e = dev.table(key ^ (0x100ull * (c + 1)), nullptr, static_cast<size_t>(pw) * ph * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
        tf.full[c] = static_cast<const float*>(ptr);
        s2.L[c] = m + static_cast<size_t>(c) * k * pw;
        s2.out[c] = static_cast<float*>(const_cast<void*>(ptr));
    }
    s2.R = chh;
    s2.M = pw;
    s2.N = ph;
    s2.K = k;
    e = launch_diffusion_atb(s2, 3, stream);
    if (e != hipSuccess) return e;
    if (dev.transfers.size() > 64) dev.transfers.clear();
    dev.transfers.emplace(key, tf);
    out.transfer = tf;
    return hipSuccess;
}

hipError_t run_diffusion(DeviceState& dev, const DiffusionStage& st, float* const planes[3], int width, int height,
                         float* padded, hipStream_t stream, const std::function<void(int)>& mark) {
    const int pw = st.pw, ph = st.ph, r = st.radius;
    const long long pplane = static_cast<long long>(pw) * ph;
    float* ext[3];
    float* tr[3];
    for (int c = 0; c < 3; ++c) {
        ext[c] = padded + c * pplane;
        tr[c] = padded + (3 + c) * pplane;
    }
    // 1) DCT-II of the plane's own rows along their reflect-101 extension to
    //    pw samples: planes (height x width) -> ext (height x pw). The padded
    //    plane's other rows are reflections of these, so they are not
    //    transformed twice.
    DctRowsArgs a{};
    hipError_t e = fft_tables(dev, pw, a);
    if (e != hipSuccess) return e;
    a.rows = height;
    a.in_stride = width;
    a.out_stride = pw;
    a.pad = r;
    a.src_w = width;
    for (int c = 0; c < 3; ++c) {
        a.in[c] = planes[c];
        a.out[c] = ext[c];
    }
    e = launch_dct2_rows(a, 3, stream);
    if (e != hipSuccess) return e;
    // 2) ext -> tr (pw x ph): column u of the padded plane, its rows gathered
    //    through the reflect-101 row extension.
    TransposeArgs t{};
    t.width = pw;
    t.height = ph;
    t.in_stride = pw;
    t.out_stride = ph;
    t.src_rows = height;
    t.row_pad = r;
    for (int c = 0; c < 3; ++c) {
        t.in[c] = ext[c];
        t.out[c] = tr[c];
    }
    e = launch_transpose(t, 3, stream);
    if (e != hipSuccess) return e;
    // 3) along padded columns: DCT-II, x T, inverse; only the frame's rows
    //    [r, r + height) are kept: tr -> ext (pw x height).
    DctRowsArgs b{};
    e = fft_tables(dev, ph, b);
    if (e != hipSuccess) return e;
    b.rows = pw;
    b.in_stride = ph;
    b.out_stride = height;
    b.out_begin = r;
    b.out_count = height;
    for (int c = 0; c < 3; ++c) {
        b.in[c] = tr[c];
        b.out[c] = ext[c];
    }
    if (mark) mark(1);
    e = launch_dct_filter_rows(b, st.transfer, 3, stream);
    if (e != hipSuccess) return e;
    if (mark) mark(2);
    // 4) ext (pw x height) -> tr (height x pw)
    t = TransposeArgs{};
    t.width = height;
    t.height = pw;
    t.in_stride = height;
    t.out_stride = pw;
    for (int c = 0; c < 3; ++c) {
        t.in[c] = ext[c];
        t.out[c] = tr[c];
    }
    e = launch_transpose(t, 3, stream);
    if (e != hipSuccess) return e;
    // 5) inverse along the frame's rows, cropped to their own samples and
    //    blended into the planes
    DctRowsArgs o = a;
    o.rows = height;
    o.in_stride = pw;
    o.out_stride = width;
    o.fraction = st.fraction;
    for (int c = 0; c < 3; ++c) {
        o.in[c] = tr[c];
        o.out[c] = planes[c];
        o.mix[c] = planes[c];
    }
    return launch_dct3_rows(o, 3, stream);
}

}  // namespace fjx
