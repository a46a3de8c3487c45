// fj_halation_host.cpp — validation of FjScatterHalation and the per-frame
// stage tables of the chunk-parallel halation kernels.
//
// Validation follows decode_optics / decode_halation_gaussian
// (native/juicer_cuda_descriptors.cpp:141-297); the stage algebra follows
// launch_scatter_halation (src/Cuda/Film/JuicerCudaScatterHalation.cu:690-852)
// with the tail amplitudes of src/ScatterHalation.h:62-64.
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "fj_halation.h"
#include "fj_halation_host.h"

namespace fjx {

namespace {

constexpr float kTailAmplitude[3] = {0.1633f, 0.6496f, 0.1870f};

bool valid_gaussian(const FjHalationGaussian& g) {
    const bool fir = g.kind == FJ_HALATION_FIR_REFLECT;
    if (g.kind > FJ_HALATION_YVV_REPLICATE_FACTORED || (fir ? g.radius < 1 || g.radius > 9 : g.radius != 0))
        return false;
    const int taps = fir ? 2 * g.radius + 1 : 0;
    for (int i = 0; i < 19; ++i) {
        const float w = g.fir_weights[i];
        if (i < taps ? !std::isfinite(w) || w < 0.0f : w != 0.0f || std::signbit(w)) return false;
    }
    if (g.kind == FJ_HALATION_YVV_REPLICATE || g.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
        for (float v : g.yvv)
            if (!std::isfinite(v)) return false;
        if (g.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
            const float h = g.yvv[0], gg = g.yvv[1], t = g.yvv[2];
            if (!(h > 0.0f && h < 1.0f && gg > 0.0f && static_cast<double>(gg) < 2.0 * (1.0 + t) && t > 0.0f &&
                  t < 1.0f) ||
                g.yvv[3] != 0.0f || std::signbit(g.yvv[3]))
                return false;
        }
    } else {
        for (float v : g.yvv)
            if (v != 0.0f || std::signbit(v)) return false;
    }
    return true;
}

using Mat3 = double[3][3];

void mat_mul3(const Mat3 a, const Mat3 b, Mat3 out) {
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    std::memcpy(out, r, sizeof(r));
}

bool solve3(Mat3 m, double b[3]) {  // Gaussian elimination with partial pivoting, in place
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::abs(m[r][c]) > std::abs(m[piv][c])) piv = r;
        if (!(std::abs(m[piv][c]) > 0.0)) return false;
        for (int k = 0; k < 3; ++k) std::swap(m[c][k], m[piv][k]);
        std::swap(b[c], b[piv]);
        for (int r = c + 1; r < 3; ++r) {
            const double f = m[r][c] / m[c][c];
            for (int k = c; k < 3; ++k) m[r][k] -= f * m[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 2; c >= 0; --c) {
        for (int k = c + 1; k < 3; ++k) b[c] -= m[c][k] * b[k];
        b[c] /= m[c][c];
    }
    return true;
}

// Factored recursion {h, g, t}, output gain and first-sample seed equivalent
// to a halation Gaussian's recursion (halation_frame.rs:85-145 derives the
// same factors from the unrounded coefficients).
struct Factored {
    double h, g, t, kappa, seed[3];
};

bool factor(const FjHalationGaussian& gs, Factored& f) {
    if (gs.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
        f = {gs.yvv[0], gs.yvv[1], gs.yvv[2], 1.0, {1.0, 1.0, 0.0}};
        return true;
    }
    // Direct form y = B x + a1 y1 + a2 y2 + a3 y3: the real pole 1 - h is the
    // root in (0, 1) of z^3 - a1 z^2 - a2 z - a3, written in h = 1 - z.
    const double B = gs.yvv[0], a1 = gs.yvv[1], a2 = gs.yvv[2], a3 = gs.yvv[3];
    const double b = 1.0 - a1 - a2 - a3;
    const double c1 = 3.0 - 2.0 * a1 - a2, c2 = 3.0 - a1;
    auto poly = [&](double h) { return ((-h + c2) * h - c1) * h + b; };
    if (!(poly(0.0) > 0.0 && poly(1.0) < 0.0)) return false;
    double lo = 0.0, hi = 1.0;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        (poly(mid) > 0.0 ? lo : hi) = mid;
    }
    const double h = 0.5 * (lo + hi), r = 1.0 - h;
    const double t = a3 / r;            // product of the other two poles
    const double g = 1.0 + t - (a1 - r);  // 1 + t - (sum of the other two poles)
    if (!(h > 0.0 && h < 1.0 && t > 0.0 && t < 1.0 && g > 0.0 && std::isfinite(g))) return false;
    f.h = h;
    f.g = g;
    f.t = t;
    f.kappa = B / (h * g);  // direct-form DC gain over the factored form's unit gain
    // Seed: the factored state whose zero-input outputs, times kappa, equal
    // the direct form's zero-input outputs from past outputs (1, 1, 1).
    Mat3 m;
    for (int col = 0; col < 3; ++col) {
        double p = col == 0, y = col == 1, d = col == 2;
        for (int n = 0; n < 3; ++n) {
            p = (1.0 - h) * p;
            d = t * d + g * (p - y);
            y = y + d;
            m[n][col] = y;
        }
    }
    double s1 = 1.0, s2 = 1.0, s3 = 1.0, zi[3];
    for (int n = 0; n < 3; ++n) {
        const double v = a1 * s1 + a2 * s2 + a3 * s3;
        s3 = s2;
        s2 = s1;
        s1 = v;
        zi[n] = v / f.kappa;
    }
    if (!solve3(m, zi)) return false;
    for (int i = 0; i < 3; ++i) f.seed[i] = zi[i];
    return std::isfinite(f.kappa) && std::isfinite(zi[0]) && std::isfinite(zi[1]) && std::isfinite(zi[2]);
}

// pw[n] = (A^L)^n for n < count, A the zero-input transition of (p, y, d):
// p' = (1-h) p; d' = t d + g (p' - y); y' = y + d'.
void chunk_powers(const Factored& f, int L, float (*pw)[9], int count) {
    const double h = f.h, g = f.g, t = f.t;
    Mat3 A = {{1.0 - h, 0.0, 0.0}, {g * (1.0 - h), 1.0 - g, t}, {g * (1.0 - h), -g, t}};
    Mat3 M = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int e = L; e > 0; e >>= 1) {
        if (e & 1) mat_mul3(M, A, M);
        mat_mul3(A, A, A);
    }
    Mat3 P = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int n = 0; n < count; ++n) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) pw[n][i * 3 + j] = static_cast<float>(P[i][j]);
        mat_mul3(P, M, P);
    }
}

bool put_gaussian(const FjHalationGaussian& g, float omega, int Lr, int Lc, HalStageDev& st) {
    HalGaussDev& out = st.g[st.count];
    std::memset(&out, 0, sizeof(out));
    out.radius = g.radius;
    std::memcpy(out.w, g.fir_weights, sizeof(out.w));
    if (g.kind == FJ_HALATION_YVV_REPLICATE || g.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
        Factored f;
        if (!factor(g, f)) return false;
        out.kind = kHalRecursive;
        out.h = static_cast<float>(f.h);
        out.g = static_cast<float>(f.g);
        out.t = static_cast<float>(f.t);
        out.kappa = static_cast<float>(f.kappa);
        for (int i = 0; i < 3; ++i) out.seed[i] = static_cast<float>(f.seed[i]);
        // The transition powers follow the f32 factors the kernel steps with.
        const Factored used{out.h, out.g, out.t, 1.0, {0, 0, 0}};
        chunk_powers(used, Lr, out.row_pow, 65);
        chunk_powers(used, Lc, out.col_pow, 5);
    } else {
        out.kind = g.kind == FJ_HALATION_FIR_REFLECT ? kHalFir : kHalIdentity;
    }
    st.omega[st.count] = omega;
    ++st.count;
    return true;
}

}  // namespace

bool halation_validate(const FjScatterHalation& h, std::string& field) {
    if (h.recipe_hash == 0) return field = "optics.scatter_halation.recipe_hash", false;
    if (!(h.scatter_amount >= 0.0f && h.scatter_amount <= 2.0f) || h.bounce_count > FJ_HALATION_MAX_BOUNCES ||
        h.renormalize > 1 || (h.bounce_count == 0 && h.renormalize))
        return field = "optics.scatter_halation.schedule", false;
    double sum = 0.0;
    for (uint32_t j = 0; j < FJ_HALATION_MAX_BOUNCES; ++j) {
        const float w = h.bounce_weights[j];
        if (j < h.bounce_count ? !(w > 0.0f && w <= 1.0f) : w != 0.0f || std::signbit(w))
            return field = "optics.scatter_halation.bounce_weights", false;
        sum += w;
    }
    if (h.bounce_count != 0 && std::abs(sum - 1.0) > 4 * std::numeric_limits<float>::epsilon())
        return field = "optics.scatter_halation.bounce_weights", false;
    bool reflection = false;
    for (const FjHalationChannel& c : h.channels) {
        if (!std::isfinite(c.total_strength) || c.total_strength < 0.0f || !(c.tail_weight >= 0.0f && c.tail_weight <= 1.0f))
            return field = "optics.scatter_halation.channel", false;
        reflection |= c.total_strength > 0.0f;
        if (!valid_gaussian(c.core)) return field = "optics.scatter_halation.gaussian", false;
        for (const FjHalationGaussian& g : c.tail)
            if (!valid_gaussian(g)) return field = "optics.scatter_halation.gaussian", false;
        for (uint32_t j = 0; j < FJ_HALATION_MAX_BOUNCES; ++j) {
            if (j >= h.bounce_count && c.bounce[j].kind != FJ_HALATION_IDENTITY)
                return field = "optics.scatter_halation.unused_bounce", false;
            if (!valid_gaussian(c.bounce[j])) return field = "optics.scatter_halation.gaussian", false;
        }
    }
    if (reflection != (h.bounce_count != 0)) return field = "optics.scatter_halation.reflection", false;
    return true;
}

bool halation_stages(const FjScatterHalation& h, int width, int height, HalationStages& out) {
    std::memset(&out, 0, sizeof(out));
    out.row_len = (width + kHalRowThreads - 1) / kHalRowThreads;
    out.col_len = (height + kHalColChunks - 1) / kHalColChunks;
    for (int ch = 0; ch < 3; ++ch) {
        const FjHalationChannel& c = h.channels[ch];
        HalStageDev& sc = out.scatter[ch];
        sc.c0 = 1.0f;
        if (h.scatter_amount != 0.0f) {
            // (1-a) x + a ((1-w_t) core + w_t sum_i A_i tail_i)
            const float a = h.scatter_amount, tw = c.tail_weight;
            sc.c0 = 1.0f - a;
            if (tw > 0.0f)
                for (int i = 0; i < 3; ++i)
                    if (!put_gaussian(c.tail[i], a * tw * kTailAmplitude[i], out.row_len, out.col_len, sc)) return false;
            if (!put_gaussian(c.core, a * (1.0f - tw), out.row_len, out.col_len, sc)) return false;
            out.scatter_active = true;
        }
        HalStageDev& bc = out.bounce[ch];
        bc.c0 = 1.0f;
        if (c.total_strength != 0.0f) {
            // (x + s sum_j b_j G_j x) / (1 + s) when renormalizing.
            const float s = c.total_strength;
            const float norm = h.renormalize ? 1.0f / (1.0f + s) : 1.0f;
            bc.c0 = norm;
            for (uint32_t j = 0; j < h.bounce_count; ++j)
                if (!put_gaussian(c.bounce[j], s * h.bounce_weights[j] * norm, out.row_len, out.col_len, bc))
                    return false;
            out.bounce_active = true;
        }
    }
    return true;
}

}  // namespace fjx
