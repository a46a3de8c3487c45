// fj_scanner_lut.cpp — native scanner PCHIP LUT construction.
//
// The reference builds this table on the Rust side and hands it back through
// the FjHostPreparation.scanner_lut callback (native/juicer_host_preparation.cpp:28-78,
// rust/film-juicer-core/src/scanner/lut.rs:140-299). This library builds it
// itself when no callback is supplied, in double precision, producing the
// same six float arrays: log2 XYZ at r^3 nodes, PCHIP slopes along C, M and Y,
// and per-cell minima/maxima used to clamp the tricubic interpolant.
#include "fj_host.h"

#include <cmath>
#include <vector>

namespace fjx {

namespace {

// Fritsch–Carlson slopes (harmonic mean of neighbouring secants, zero at
// extrema) with the three-point one-sided endpoint rule.
void line_slopes(const double* v, int n, double* out) {
    std::vector<double> d(static_cast<size_t>(n - 1));
    for (int i = 0; i + 1 < n; ++i) d[i] = v[i + 1] - v[i];
    if (n == 2) {
        out[0] = out[1] = d[0];
        return;
    }
    auto end_slope = [](double a, double b) {
        const double s = 0.5 * (3.0 * a - b);
        if (s * a <= 0.0) return 0.0;
        if (a * b < 0.0 && std::fabs(s) > std::fabs(3.0 * a)) return 3.0 * a;
        return s;
    };
    out[0] = end_slope(d[0], d[1]);
    out[n - 1] = end_slope(d[n - 2], d[n - 3]);
    for (int i = 1; i + 1 < n; ++i) {
        const double p = d[i - 1], q = d[i];
        out[i] = (p == 0.0 || q == 0.0 || p * q <= 0.0) ? 0.0 : 2.0 * p * q / (p + q);
    }
}

}  // namespace

bool build_scanner_lut(const ScannerLutInput& in, ScannerLut& out, std::string& diagnostic, uint64_t* fail_node,
                       uint32_t* fail_channel) {
    const int r = in.resolution;
    if (r < 2 || r > 256) {
        diagnostic = "scanner LUT resolution must be in [2, 256]";
        return false;
    }
    const int k = in.sample_count;
    const size_t count = static_cast<size_t>(r) * r * r * 3;
    std::vector<double> nodes(count);
    auto at = [r](int c, int m, int y, int ch) { return ((static_cast<size_t>(c) * r + m) * r + y) * 3 + ch; };
    const float* weights[3] = {in.weighted_x, in.weighted_y, in.weighted_z};
    for (int c = 0; c < r; ++c) {
        for (int m = 0; m < r; ++m) {
            for (int y = 0; y < r; ++y) {
                const double nrm[3] = {double(c) / (r - 1), double(m) / (r - 1), double(y) / (r - 1)};
                double dens[3];
                for (int i = 0; i < 3; ++i) {
                    const double scaled = nrm[i] / static_cast<double>(in.inverse_span_cmy[i]);
                    dens[i] = in.medium == 0 ? scaled - in.range_min_cmy[i] : scaled + in.range_min_cmy[i];
                }
                double xyz[3] = {0.0, 0.0, 0.0};
                if (k <= 0) xyz[0] = xyz[1] = xyz[2] = NAN;
                for (int s = 0; s < k; ++s) {
                    const double base = in.baseline ? static_cast<double>(in.baseline[s]) : 0.0;
                    const double od = dens[0] * in.eps_c[s] + dens[1] * in.eps_m[s] + dens[2] * in.eps_y[s] + base;
                    const double t = std::pow(10.0, -od);
                    for (int ch = 0; ch < 3; ++ch) {
                        const double w = weights[ch][s];
                        if (!std::isfinite(w)) continue;
                        const double prod = t * w;
                        if (!std::isnan(prod)) xyz[ch] += prod;
                    }
                }
                for (int ch = 0; ch < 3; ++ch) {
                    const double y_norm = xyz[ch] * static_cast<double>(in.inv_yn);
                    const double pos = std::isnan(y_norm) ? 0.0 : std::fmax(y_norm, 0.0);
                    const double l10 = k > 0 ? std::log10(pos + 1e-10) : NAN;
                    if (!std::isfinite(l10)) {
                        diagnostic = "scanner LUT build produced a non-finite log2 XYZ node";
                        if (fail_node) *fail_node = (static_cast<uint64_t>(c) * r + m) * r + y;
                        if (fail_channel) *fail_channel = static_cast<uint32_t>(ch);
                        return false;
                    }
                    nodes[at(c, m, y, ch)] = l10 * 3.32192809488736234787;
                }
            }
        }
    }
    out.resolution = r;
    out.log2_xyz.resize(count);
    out.slope_c.resize(count);
    out.slope_m.resize(count);
    out.slope_y.resize(count);
    for (size_t i = 0; i < count; ++i) out.log2_xyz[i] = static_cast<float>(nodes[i]);
    std::vector<double> line(r), slope(r);
    std::vector<float>* dst[3] = {&out.slope_c, &out.slope_m, &out.slope_y};
    for (int axis = 0; axis < 3; ++axis) {
        for (int a = 0; a < r; ++a) {
            for (int b = 0; b < r; ++b) {
                for (int ch = 0; ch < 3; ++ch) {
                    auto idx = [&](int i) {
                        return axis == 0 ? at(i, a, b, ch) : axis == 1 ? at(a, i, b, ch) : at(a, b, i, ch);
                    };
                    for (int i = 0; i < r; ++i) line[i] = nodes[idx(i)];
                    line_slopes(line.data(), r, slope.data());
                    for (int i = 0; i < r; ++i) (*dst[axis])[idx(i)] = static_cast<float>(slope[i]);
                }
            }
        }
    }
    const int s = r - 1;
    const size_t cells = static_cast<size_t>(s) * s * s * 3;
    out.cell_min.resize(cells);
    out.cell_max.resize(cells);
    for (int c = 0; c < s; ++c) {
        for (int m = 0; m < s; ++m) {
            for (int y = 0; y < s; ++y) {
                for (int ch = 0; ch < 3; ++ch) {
                    double lo = nodes[at(c, m, y, ch)], hi = lo;
                    for (int corner = 1; corner < 8; ++corner) {
                        const double v = nodes[at(c + (corner >> 2), m + ((corner >> 1) & 1), y + (corner & 1), ch)];
                        lo = v < lo ? v : lo;
                        hi = hi < v ? v : hi;
                    }
                    const size_t i = ((static_cast<size_t>(c) * s + m) * s + y) * 3 + ch;
                    out.cell_min[i] = static_cast<float>(lo);
                    out.cell_max[i] = static_cast<float>(hi);
                }
            }
        }
    }
    return true;
}

}  // namespace fjx
