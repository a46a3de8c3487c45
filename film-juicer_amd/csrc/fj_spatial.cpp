// fj_spatial.cpp — host side of the DCT-domain spatial operators.
//
// The spatial DIR filter of the reference is a weighted sum of separable
// Gaussian-like components, each run with half-sample reflected borders:
// short ones as an FIR (JuicerCudaFilmPipeline.cu:388-436), long ones as a
// third-order Young–van Vliet recursion whose start and end states are solved
// for the 2N-periodic reflected signal (rust film/dir/boundary.rs). Both are
// exact convolutions on the half-sample symmetric periodic extension, which
// the DCT-II diagonalizes; this file builds their per-axis eigenvalues and
// drives the five-kernel transform (row DCT, transpose, column DCT x T x
// column IDCT, transpose, row IDCT).
#include "fj_spatial.h"

#include <cmath>
#include <cstring>
#include <vector>

#include "fj_host.h"

namespace fjx {

std::vector<float> gaussian_weights(int radius, float sigma_px) {
    // Restates the prepared Gaussian kernel (rust optics/gaussian.rs:20-56):
    // double samples exp(-x^2 / 2 sigma^2), each rounded to float, then
    // normalized by the unrounded double sum.
    std::vector<float> w;
    if (radius < 1 || !(sigma_px > 0.0f) || !std::isfinite(sigma_px)) return w;
    const double s2 = double(sigma_px) * double(sigma_px) * 2.0;
    double sum = 0.0;
    for (int i = -radius; i <= radius; ++i) {
        const double x = i;
        const double v = std::exp(-(x * x) / s2);
        w.push_back(static_cast<float>(v));
        sum += v;
    }
    const double inv = sum != 0.0 ? 1.0 / sum : 0.0;
    for (float& v : w) v = static_cast<float>(double(v) * inv);
    return w;
}

void fir_eigenvalues(const std::vector<float>& w, int n, float* out) {
    const int r = static_cast<int>(w.size() / 2);
    const double pi = 3.14159265358979323846264338327950;
    for (int k = 0; k < n; ++k) {
        const double theta = pi * k / n;
        double t = 0.0;
        for (int j = -r; j <= r; ++j) t += double(w[j + r]) * std::cos(theta * j);
        out[k] = static_cast<float>(t);
    }
}

void yvv_eigenvalues(double b, const double a[3], int n, float* out) {
    // Forward then backward pass of y[i] = b x[i] + a1 y[i-1] + a2 y[i-2] + a3 y[i-3]:
    // |H|^2 = b^2 / |1 - sum a_m e^{-i m theta}|^2.
    const double pi = 3.14159265358979323846264338327950;
    for (int k = 0; k < n; ++k) {
        const double theta = pi * k / n;
        double re = 1.0, im = 0.0;
        for (int m = 1; m <= 3; ++m) {
            re -= a[m - 1] * std::cos(m * theta);
            im += a[m - 1] * std::sin(m * theta);
        }
        out[k] = static_cast<float>(b * b / (re * re + im * im));
    }
}

bool dir_transfer(const FjSpatialDir& sd, int width, int height, HostTransfer& out, std::string& diag) {
    out = HostTransfer{};
    if (sd.component_count < 1 || sd.component_count > 4) {
        diag = "spatial_dir.component_count";
        return false;
    }
    for (size_t i = 0; i < sd.component_count; ++i) {
        const FjDirGaussian& g = sd.components[i];
        if (!std::isfinite(g.weight) || g.weight < 0.0f) {
            diag = "spatial_dir.component.weight";
            return false;
        }
        if (!(g.weight > 0.0f)) continue;
        std::vector<float> tx(width), ty(height);
        switch (g.reference_operator) {
            case FJ_DIR_OPERATOR_IDENTITY:
                std::fill(tx.begin(), tx.end(), 1.0f);
                std::fill(ty.begin(), ty.end(), 1.0f);
                break;
            case FJ_DIR_OPERATOR_FIR_REFLECT: {
                if (!(g.sigma_px > 0.0f && g.sigma_px < 3.0f) || g.radius <= 0) {
                    diag = "spatial_dir.component.fir";
                    return false;
                }
                const std::vector<float> w = kernel_weights(g.radius, g.sigma_px);
                fir_eigenvalues(w, width, tx.data());
                fir_eigenvalues(w, height, ty.data());
                break;
            }
            case FJ_DIR_OPERATOR_YVV_REFLECT: {
                const double dc = g.feedforward + g.feedback[0] + g.feedback[1] + g.feedback[2];
                if (!(g.sigma_px >= 3.0f) || !std::isfinite(g.feedforward) || !(g.feedforward > 0.0) ||
                    !std::isfinite(dc) || std::fabs(dc - 1.0) > 1.0e-12) {
                    diag = "spatial_dir.component.yvv";
                    return false;
                }
                yvv_eigenvalues(g.feedforward, g.feedback, width, tx.data());
                yvv_eigenvalues(g.feedforward, g.feedback, height, ty.data());
                break;
            }
            default:
                diag = "spatial_dir.component.operator";
                return false;
        }
        out.weights.push_back(g.weight);
        out.rows.push_back(std::move(tx));   // first axis: along x
        out.cols.push_back(std::move(ty));   // second axis: along y
    }
    if (out.weights.empty()) {
        diag = "spatial_dir.no_active_component";
        return false;
    }
    return true;
}

hipError_t run_dct_filter(DeviceState& dev, const DctFilterJob& job, hipStream_t stream) {
    const int W = job.width, H = job.height;
    DctRowsArgs a{};
    const void* ptr = nullptr;
    hipError_t e = fft_tables(dev, W, a);
    if (e != hipSuccess) return e;
    // 1) DCT-II along rows: in -> tmp_a
    a.rows = H;
    a.in_stride = W;
    a.out_stride = W;
    for (int c = 0; c < job.planes; ++c) {
        a.in[c] = job.in[c];
        a.out[c] = job.tmp_a[c];
    }
    e = launch_dct2_rows(a, job.planes, stream);
    if (e != hipSuccess) return e;
    // 2) transpose tmp_a (H x W) -> tmp_b (W x H)
    TransposeArgs t{};
    t.width = W;
    t.height = H;
    t.in_stride = W;
    t.out_stride = H;
    for (int c = 0; c < job.planes; ++c) {
        t.in[c] = job.tmp_a[c];
        t.out[c] = job.tmp_b[c];
    }
    e = launch_transpose(t, job.planes, stream);
    if (e != hipSuccess) return e;
    // 3) along columns: DCT-II, x T, inverse; in place on tmp_b
    DctRowsArgs b{};
    e = fft_tables(dev, H, b);
    if (e != hipSuccess) return e;
    b.rows = W;
    b.in_stride = H;
    b.out_stride = H;
    for (int c = 0; c < job.planes; ++c) {
        b.in[c] = job.tmp_b[c];
        b.out[c] = job.tmp_b[c];
    }
    e = launch_dct_filter_rows(b, job.transfer, job.planes, stream);
    if (e != hipSuccess) return e;
    // 4) transpose back tmp_b (W x H) -> tmp_a (H x W)
    t.width = H;
    t.height = W;
    t.in_stride = H;
    t.out_stride = W;
    for (int c = 0; c < job.planes; ++c) {
        t.in[c] = job.tmp_b[c];
        t.out[c] = job.tmp_a[c];
    }
    e = launch_transpose(t, job.planes, stream);
    if (e != hipSuccess) return e;
    // 5) inverse along rows: tmp_a -> out
    for (int c = 0; c < job.planes; ++c) {
        a.in[c] = job.tmp_a[c];
        a.out[c] = job.out[c];
    }
    (void)ptr;
    return launch_dct3_rows(a, job.planes, stream);
}

hipError_t fft_tables(DeviceState& dev, int n, DctRowsArgs& a) {
    if (!make_fft_plan(n, a.plan)) return hipErrorInvalidValue;
    const void* ptr = nullptr;
    const uint64_t key_tw = 0xFF7A000000000000ull ^ static_cast<uint64_t>(n);
    const uint64_t key_half = 0xFF7B000000000000ull ^ static_cast<uint64_t>(n);
    if (dev.tables.count(key_tw) == 0) {
        const std::vector<float> tw = fft_twiddles(n);
        hipError_t e = dev.table(key_tw, tw.data(), tw.size() * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
    }
    hipError_t e = dev.table(key_tw, nullptr, 0, &ptr);
    if (e != hipSuccess) return e;
    a.table = static_cast<const float2*>(ptr);
    if (dev.tables.count(key_half) == 0) {
        const std::vector<float> hw = dct_half_twiddles(n);
        e = dev.table(key_half, hw.data(), hw.size() * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
    }
    e = dev.table(key_half, nullptr, 0, &ptr);
    if (e != hipSuccess) return e;
    a.half_tw = static_cast<const float2*>(ptr);
    return hipSuccess;
}

hipError_t upload_transfer(DeviceState& dev, const HostTransfer& h, uint64_t key, DctTransfer& out) {
    out = DctTransfer{};
    out.count = static_cast<int>(h.weights.size());
    for (int c = 0; c < out.count; ++c) {
        const void* ptr = nullptr;
        hipError_t e = dev.table(key ^ (0x100ull * (2 * c + 1)), h.rows[c].data(), h.rows[c].size() * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
        out.t_rows[c] = static_cast<const float*>(ptr);
        e = dev.table(key ^ (0x100ull * (2 * c + 2)), h.cols[c].data(), h.cols[c].size() * sizeof(float), &ptr);
        if (e != hipSuccess) return e;
        out.t_cols[c] = static_cast<const float*>(ptr);
        out.weight[c] = h.weights[c];
    }
    return hipSuccess;
}

}  // namespace fjx
