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

// State transition of one recursion step with zero input, in the state
// order the kernels keep.
void transition(const FjHalationGaussian& g, Mat3 A) {
    const double c0 = g.yvv[0], c1 = g.yvv[1], c2 = g.yvv[2], c3 = g.yvv[3];
    if (g.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
        // p' = (1-h) p; d' = t d + g (p' - y); y' = y + d'.
        const double h = c0, gg = c1, t = c2;
        const double a[3][3] = {{1.0 - h, 0.0, 0.0}, {gg * (1.0 - h), 1.0 - gg, t}, {gg * (1.0 - h), -gg, t}};
        std::memcpy(A, a, sizeof(a));
    } else {
        const double a[3][3] = {{c1, c2, c3}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}};
        std::memcpy(A, a, sizeof(a));
    }
}

// pw[n] = (A^L)^n for n < count.
void chunk_powers(const FjHalationGaussian& g, int L, double (*pw)[9], int count) {
    Mat3 A, M = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    transition(g, A);
    for (int e = L; e > 0; e >>= 1) {
        if (e & 1) mat_mul3(M, A, M);
        mat_mul3(A, A, A);
    }
    Mat3 P = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int n = 0; n < count; ++n) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) pw[n][i * 3 + j] = P[i][j];
        mat_mul3(P, M, P);
    }
}

void put_gaussian(const FjHalationGaussian& g, float omega, int Lr, int Lc, HalStageDev& st) {
    HalGaussDev& out = st.g[st.count];
    std::memset(&out, 0, sizeof(out));
    out.kind = static_cast<int>(g.kind);
    out.radius = g.radius;
    std::memcpy(out.w, g.fir_weights, sizeof(out.w));
    for (int i = 0; i < 4; ++i) out.c[i] = g.yvv[i];
    if (g.kind == FJ_HALATION_YVV_REPLICATE || g.kind == FJ_HALATION_YVV_REPLICATE_FACTORED) {
        chunk_powers(g, Lr, out.row_pow, 65);
        chunk_powers(g, Lc, out.col_pow, 5);
    }
    st.omega[st.count] = omega;
    ++st.count;
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

void halation_stages(const FjScatterHalation& h, int width, int height, HalationStages& out) {
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
                for (int i = 0; i < 3; ++i) put_gaussian(c.tail[i], a * tw * kTailAmplitude[i], out.row_len, out.col_len, sc);
            put_gaussian(c.core, a * (1.0f - tw), out.row_len, out.col_len, sc);
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
                put_gaussian(c.bounce[j], s * h.bounce_weights[j] * norm, out.row_len, out.col_len, bc);
            out.bounce_active = true;
        }
    }
}

}  // namespace fjx
