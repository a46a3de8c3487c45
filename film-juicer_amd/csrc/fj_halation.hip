// fj_halation.hip — chunk-parallel scatter/halation stages for gfx950.
//
// Reference operator: src/Cuda/Film/JuicerCudaScatterHalation.cu
//   reflect_index :30-46, finalize_scatter_value :57-69, FIR :104-180,
//   yvv_step :183-198, YVV row/column passes :200-420 (rows left->right then
//   right->left; columns bottom->top then top->bottom; every pass seeded with
//   its first sample), back_reflection_finalize :443-463.
// The reference runs one thread per line with a serial recursion; here every
// line is split into chunks owned by separate threads and the chunk states are
// combined with a scan, so a UHD plane keeps the whole chip busy. Stages made
// only of FIR taps run as one tiled launch with both passes in LDS.
#include <hip/hip_runtime.h>

#include "fj_batch.h"
#include "fj_halation.h"

namespace fjx {

constexpr int kFirTaps = 19;   // FIR radius <= 9; weights are zero-padded to 19 taps
constexpr int kRowGroup = 8;   // samples loaded ahead of the serial steps (LDS)
constexpr int kColGroup = 16;  // same for global memory
constexpr int kFinGroup = 8;   // rows per step of the final weighted sweep

__device__ __forceinline__ int hal_reflect(int i, int n) {
    if (i >= 0 && i < n) return i;
    if (i >= -n && i < 0) return -i - 1;
    if (i >= n && i < 2 * n) return 2 * n - 1 - i;
    const int period = 2 * n;
    i %= period;
    if (i < 0) i += period;
    return i >= n ? period - 1 - i : i;
}

// Factored recursion (JuicerCudaScatterHalation.cu:186-191): real pole p,
// second-order section y advanced by the increment d; output kappa * y.
struct Rec {
    float h, g, t, kappa, seed[3];
};

__device__ __forceinline__ Rec load_rec(const HalGaussDev* g) {
    return Rec{g->h, g->g, g->t, g->kappa, {g->seed[0], g->seed[1], g->seed[2]}};
}

__device__ __forceinline__ float hal_step(const Rec& r, float x, float s[3]) {
    s[0] = fmaf(r.h, x - s[0], s[0]);
    s[2] = fmaf(r.t, s[2], r.g * (s[0] - s[1]));
    s[1] += s[2];
    return r.kappa * s[1];
}

__device__ __forceinline__ void hal_seed(const Rec& r, float first, float s[3]) {
    s[0] = r.seed[0] * first;
    s[1] = r.seed[1] * first;
    s[2] = r.seed[2] * first;
}

__device__ __forceinline__ void hal_zero(float s[3]) { s[0] = s[1] = s[2] = 0.0f; }

// Runs the recursion over samples i < len at load(i); with kWrite the outputs
// go to store(i, v). Each group of kG samples is loaded before its serial
// steps, so the load latency is paid once per group.
template <int kG, bool kWrite, typename Load, typename Store>
__device__ __forceinline__ void run_chunk(const Rec& r, float s[3], int len, Load load, Store store) {
    for (int b = 0; b < len; b += kG) {
        float v[kG];
#pragma unroll
        for (int i = 0; i < kG; ++i) v[i] = b + i < len ? load(b + i) : 0.0f;
#pragma unroll
        for (int i = 0; i < kG; ++i) {
            if (b + i < len) {
                const float o = hal_step(r, v[i], s);
                if constexpr (kWrite) store(b + i, o);
            }
        }
    }
}

// Exclusive scan of chunk end states E_r = e_r + M E_{r-1}: thread layout
// tid = r * kLines + line, so a wave holds kPer = 64 / kLines consecutive
// chunks of kLines lines. tab[n] = M^n (n <= kPer). Inside a wave the scan
// uses lane shuffles; the waves' totals meet in LDS. Returns E_{r-1} (zero
// for r = 0). Every thread of the block must call it.
template <int kLines>
__device__ __forceinline__ void chunk_scan(float* buf, int line, const float e[3], const float* tab,
                                           float prev[3]) {
    constexpr int kPer = 64 / kLines;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int rw = lane / kLines;
    float E0 = e[0], E1 = e[1], E2 = e[2];
#pragma unroll
    for (int d = 1; d < kPer; d <<= 1) {
        const float o0 = __shfl_up(E0, d * kLines, 64);
        const float o1 = __shfl_up(E1, d * kLines, 64);
        const float o2 = __shfl_up(E2, d * kLines, 64);
        if (rw >= d) {
            const float* P = tab + d * 9;
            E0 += fmaf(P[0], o0, fmaf(P[1], o1, P[2] * o2));
            E1 += fmaf(P[3], o0, fmaf(P[4], o1, P[5] * o2));
            E2 += fmaf(P[6], o0, fmaf(P[7], o1, P[8] * o2));
        }
    }
    float x0 = __shfl_up(E0, kLines, 64);
    float x1 = __shfl_up(E1, kLines, 64);
    float x2 = __shfl_up(E2, kLines, 64);
    if (rw == 0) x0 = x1 = x2 = 0.0f;
    if (rw == kPer - 1) {
        float* me = buf + (wave * kLines + line) * 3;
        me[0] = E0;
        me[1] = E1;
        me[2] = E2;
    }
    __syncthreads();
    // End state of the previous waves' last chunk.
    float T0 = 0.0, T1 = 0.0, T2 = 0.0f;
    const float* Mw = tab + kPer * 9;
    for (int w = 0; w < wave; ++w) {
        const float* o = buf + (w * kLines + line) * 3;
        const float n0 = fmaf(Mw[0], T0, fmaf(Mw[1], T1, fmaf(Mw[2], T2, o[0])));
        const float n1 = fmaf(Mw[3], T0, fmaf(Mw[4], T1, fmaf(Mw[5], T2, o[1])));
        const float n2 = fmaf(Mw[6], T0, fmaf(Mw[7], T1, fmaf(Mw[8], T2, o[2])));
        T0 = n0;
        T1 = n1;
        T2 = n2;
    }
    const float* P = tab + rw * 9;
    prev[0] = fmaf(P[0], T0, fmaf(P[1], T1, fmaf(P[2], T2, x0)));
    prev[1] = fmaf(P[3], T0, fmaf(P[4], T1, fmaf(P[5], T2, x1)));
    prev[2] = fmaf(P[6], T0, fmaf(P[7], T1, fmaf(P[8], T2, x2)));
    __syncthreads();
}

// chunk_scan for a pass that processes chunks in reverse thread order: the
// thread of virtual position r' = (blockDim.x / kLines - 1) - r takes the
// place of position r' in chunk_scan, so a chunk gets the same arithmetic as
// there. Every thread of the block must call it.
template <int kLines>
__device__ __forceinline__ void chunk_scan_rev(float* buf, int line, const float e[3], const float* tab,
                                               float prev[3]) {
    constexpr int kPer = 64 / kLines;
    const int lane = threadIdx.x & 63;
    const int wave = (blockDim.x >> 6) - 1 - (threadIdx.x >> 6);
    const int rw = kPer - 1 - lane / kLines;
    float E0 = e[0], E1 = e[1], E2 = e[2];
#pragma unroll
    for (int d = 1; d < kPer; d <<= 1) {
        const float o0 = __shfl_down(E0, d * kLines, 64);
        const float o1 = __shfl_down(E1, d * kLines, 64);
        const float o2 = __shfl_down(E2, d * kLines, 64);
        if (rw >= d) {
            const float* P = tab + d * 9;
            E0 += fmaf(P[0], o0, fmaf(P[1], o1, P[2] * o2));
            E1 += fmaf(P[3], o0, fmaf(P[4], o1, P[5] * o2));
            E2 += fmaf(P[6], o0, fmaf(P[7], o1, P[8] * o2));
        }
    }
    float x0 = __shfl_down(E0, kLines, 64);
    float x1 = __shfl_down(E1, kLines, 64);
    float x2 = __shfl_down(E2, kLines, 64);
    if (rw == 0) x0 = x1 = x2 = 0.0f;
    if (rw == kPer - 1) {
        float* me = buf + (wave * kLines + line) * 3;
        me[0] = E0;
        me[1] = E1;
        me[2] = E2;
    }
    __syncthreads();
    float T0 = 0.0, T1 = 0.0, T2 = 0.0f;
    const float* Mw = tab + kPer * 9;
    for (int w = 0; w < wave; ++w) {
        const float* o = buf + (w * kLines + line) * 3;
        const float n0 = fmaf(Mw[0], T0, fmaf(Mw[1], T1, fmaf(Mw[2], T2, o[0])));
        const float n1 = fmaf(Mw[3], T0, fmaf(Mw[4], T1, fmaf(Mw[5], T2, o[1])));
        const float n2 = fmaf(Mw[6], T0, fmaf(Mw[7], T1, fmaf(Mw[8], T2, o[2])));
        T0 = n0;
        T1 = n1;
        T2 = n2;
    }
    const float* P = tab + rw * 9;
    prev[0] = fmaf(P[0], T0, fmaf(P[1], T1, fmaf(P[2], T2, x0)));
    prev[1] = fmaf(P[3], T0, fmaf(P[4], T1, fmaf(P[5], T2, x1)));
    prev[2] = fmaf(P[6], T0, fmaf(P[7], T1, fmaf(P[8], T2, x2)));
    __syncthreads();
}

// FIR weights zero-padded to 19 taps centred at index 9.
__device__ __forceinline__ void padded_taps(const HalGaussDev* g, float w[kFirTaps]) {
    const int r = g->radius;
#pragma unroll
    for (int i = 0; i < kFirTaps; ++i) {
        const int o = i - 9;
        w[i] = (o >= -r && o <= r) ? g->w[o + r] : 0.0f;
    }
}

// Row pass: grid (height, 3 channels), 256 threads. For every non-identity
// Gaussian k of the channel's stage, writes the row-filtered row into q[k].
// Chunk q covers [q L, q L + L); the forward pass gives chunk t to thread t,
// the backward pass chunk 255 - t, so the processing order is the thread
// order in both.
__global__ __launch_bounds__(kHalRowThreads) void hal_rows_kernel(HalArgs a) {
    const int y = blockIdx.x;
    const int ch = blockIdx.y;
    const HalStageDev* st = a.stage + ch;
    const int count = st->count;
    if (count == 0) return;
    extern __shared__ float hal_lds[];
    __shared__ float scan[(kHalRowThreads / 64) * 3];
    const int W = a.width;
    float* xs = hal_lds;
    float* ys = hal_lds + W;
    const int t = threadIdx.x;
    const float* X = a.in[ch] + static_cast<size_t>(y) * W;
    batched<8, float>(t, W, kHalRowThreads, [&](int i) { return X[i]; }, [&](int i, float v) { xs[i] = v; });
    __syncthreads();
    const int L = a.row_len;
    const int lo = t * L;
    const int len = max(0, min(W, lo + L) - lo);
    const int qb = kHalRowThreads - 1 - t;  // backward chunk
    const int lob = qb * L;
    const int lenb = max(0, min(W, lob + L) - lob);
    const int last = (W + L - 1) / L - 1;
    auto fwd_x = [&](int i) { return xs[lo + i]; };
    auto fwd_y = [&](int i, float v) { ys[lo + i] = v; };
    auto bwd_y = [&](int i) { return ys[lob + lenb - 1 - i]; };
    auto bwd_store = [&](int i, float v) { ys[lob + lenb - 1 - i] = v; };
    auto none = [](int, float) {};
    for (int k = 0; k < count; ++k) {
        const HalGaussDev* g = st->g + k;
        const int kind = g->kind;
        if (kind == kHalIdentity) continue;
        if (kind == kHalFir) {
            float w[kFirTaps];
            padded_taps(g, w);
            for (int i = t; i < W; i += kHalRowThreads) {
                float sum = 0.0f;
#pragma unroll
                for (int o = 0; o < kFirTaps; ++o) sum = fmaf(xs[hal_reflect(i + o - 9, W)], w[o], sum);
                ys[i] = sum;
            }
        } else {
            const Rec rec = load_rec(g);
            const float* tab = &g->row_pow[0][0];
            float s[3], prev[3];
            // Left to right; chunk 0 is seeded with x[0].
            if (t == 0) hal_seed(rec, xs[0], s);
            else hal_zero(s);
            run_chunk<kRowGroup, false>(rec, s, len, fwd_x, none);
            chunk_scan<1>(scan, 0, s, tab, prev);
            if (t == 0) hal_seed(rec, xs[0], s);
            else s[0] = prev[0], s[1] = prev[1], s[2] = prev[2];
            run_chunk<kRowGroup, true>(rec, s, len, fwd_x, fwd_y);
            __syncthreads();
            // Right to left; the last (shortest) chunk is seeded with the
            // forward pass's terminal value.
            const float terminal = ys[W - 1];
            if (qb == last) hal_seed(rec, terminal, s);
            else hal_zero(s);
            run_chunk<kRowGroup, false>(rec, s, lenb, bwd_y, none);
            chunk_scan<1>(scan, 0, s, tab, prev);
            if (qb == last) hal_seed(rec, terminal, s);
            else s[0] = prev[0], s[1] = prev[1], s[2] = prev[2];
            run_chunk<kRowGroup, true>(rec, s, lenb, bwd_y, bwd_store);
        }
        __syncthreads();
        float* Q = a.q[ch][k] + static_cast<size_t>(y) * W;
        for (int i = t; i < W; i += kHalRowThreads) Q[i] = ys[i];
        __syncthreads();
    }
}

// Column pass + combination: grid (ceil(width/16), 3), 1024 threads =
// 16 columns x 64 chunks. Chunks are aligned to the bottom row, so the top
// chunk is the short one: last for the upward pass, first (seeded) for the
// downward pass. Thread j owns chunk j going up and chunk 63 - j going down
// (and in the final sweep), so the processing order is the thread order.
__global__ __launch_bounds__(kHalColWidth * kHalColChunks) void hal_cols_kernel(HalArgs a) {
    const int ch = blockIdx.y;
    const HalStageDev* st = a.stage + ch;
    const int count = st->count;
    if (count == 0) return;
    __shared__ float scan[(kHalColChunks * kHalColWidth / 64) * kHalColWidth * 3];
    const int cx = threadIdx.x % kHalColWidth;
    const int j = threadIdx.x / kHalColWidth;
    const int W = a.width, H = a.height, L = a.col_len;
    const int x = blockIdx.x * kHalColWidth + cx;
    const bool live = x < W;
    const int xc = live ? x : W - 1;
    const int hi = H - 1 - j * L;  // upward chunk j: rows hi .. lo
    const int lo = max(0, H - (j + 1) * L);
    const int len = max(0, hi - lo + 1);
    const int jd = kHalColChunks - 1 - j;  // downward chunk
    const int hid = H - 1 - jd * L;
    const int lod = max(0, H - (jd + 1) * L);
    const int lend = max(0, hid - lod + 1);
    const int top = (H + L - 1) / L - 1;  // chunk holding row 0
    const float* Xin = a.in[ch];
    float* Xout = a.out[ch];
    float S[kHalMaxGauss][3];
#pragma unroll
    for (int k = 0; k < kHalMaxGauss; ++k) {
        hal_zero(S[k]);
        if (k >= count) continue;
        const HalGaussDev* g = st->g + k;
        if (g->kind != kHalRecursive) continue;
        const Rec rec = load_rec(g);
        const float* tab = &g->col_pow[0][0];
        float* Q = a.q[ch][k];
        auto up = [&](int i) { return Q[static_cast<size_t>(hi - i) * W + xc]; };
        auto up_store = [&](int i, float v) {
            if (live) Q[static_cast<size_t>(hi - i) * W + xc] = v;
        };
        auto down = [&](int i) { return Q[static_cast<size_t>(lod + i) * W + xc]; };
        auto none = [](int, float) {};
        float s[3], prev[3];
        // Bottom to top, seeded with the bottom row.
        const float bottom = Q[static_cast<size_t>(H - 1) * W + xc];
        if (j == 0) hal_seed(rec, bottom, s);
        else hal_zero(s);
        run_chunk<kColGroup, false>(rec, s, len, up, none);
        chunk_scan<kHalColWidth>(scan, cx, s, tab, prev);
        if (j == 0) hal_seed(rec, bottom, s);
        else s[0] = prev[0], s[1] = prev[1], s[2] = prev[2];
        run_chunk<kColGroup, true>(rec, s, len, up, up_store);
        __syncthreads();  // the downward chunks read other threads' upward output
        // Top to bottom, seeded with row 0 of the upward result.
        const float first = Q[xc];
        if (jd == top) hal_seed(rec, first, s);
        else hal_zero(s);
        run_chunk<kColGroup, false>(rec, s, lend, down, none);
        chunk_scan<kHalColWidth>(scan, cx, s, tab, prev);
        if (jd == top) hal_seed(rec, first, S[k]);
        else S[k][0] = prev[0], S[k][1] = prev[1], S[k][2] = prev[2];
    }
    // Final downward sweep: every Gaussian's column output, weighted sum.
    const float c0 = st->c0;
    for (int b = 0; b < lend; b += kFinGroup) {
        float acc[kFinGroup], xv[kFinGroup];
#pragma unroll
        for (int i = 0; i < kFinGroup; ++i) {
            xv[i] = b + i < lend ? Xin[static_cast<size_t>(lod + b + i) * W + xc] : 0.0f;
            acc[i] = c0 * xv[i];
        }
#pragma unroll
        for (int k = 0; k < kHalMaxGauss; ++k) {
            if (k >= count) continue;
            const HalGaussDev* g = st->g + k;
            const float om = st->omega[k];
            const int kind = g->kind;
            if (kind == kHalIdentity) {
#pragma unroll
                for (int i = 0; i < kFinGroup; ++i) acc[i] = fmaf(om, xv[i], acc[i]);
            } else if (kind == kHalFir) {
                const float* Q = a.q[ch][k];
                float w[kFirTaps];
                padded_taps(g, w);
                float win[kFinGroup + kFirTaps - 1];
#pragma unroll
                for (int i = 0; i < kFinGroup + kFirTaps - 1; ++i)
                    win[i] = Q[static_cast<size_t>(hal_reflect(lod + b + i - 9, H)) * W + xc];
#pragma unroll
                for (int i = 0; i < kFinGroup; ++i) {
                    float sum = 0.0f;
#pragma unroll
                    for (int o = 0; o < kFirTaps; ++o) sum = fmaf(win[i + o], w[o], sum);
                    acc[i] = fmaf(om, sum, acc[i]);
                }
            } else {
                const Rec rec = load_rec(g);
                const float* Q = a.q[ch][k];
                float qv[kFinGroup];
#pragma unroll
                for (int i = 0; i < kFinGroup; ++i)
                    qv[i] = b + i < lend ? Q[static_cast<size_t>(lod + b + i) * W + xc] : 0.0f;
#pragma unroll
                for (int i = 0; i < kFinGroup; ++i)
                    if (b + i < lend) acc[i] = fmaf(om, hal_step(rec, qv[i], S[k]), acc[i]);
            }
        }
        if (live) {
#pragma unroll
            for (int i = 0; i < kFinGroup; ++i)
                if (b + i < lend) Xout[static_cast<size_t>(lod + b + i) * W + xc] = acc[i];
        }
    }
}

// Column pass + combination for columns of at most 64 * kHalColRegRows rows:
// the layout of hal_cols_kernel, but thread j owns chunk j in both
// directions and holds it in registers (v: the Gaussian's rows, acc: the
// weighted sum), so Q_k is read once and never written back. Slot m of a
// thread holds row hi - (R - 1) + m; the chunk's rows are the slots
// m >= m0 = R - len. Slots below m0 load some row of the plane (clamped) and
// their results are discarded, so no load or step needs a branch: the upward
// pass (slots R-1 .. 0) takes its end state and its last output at slot m0,
// the downward pass (0 .. R-1) resets its state at slot m0. The downward
// scan is chunk_scan_rev, which gives every chunk the arithmetic of
// hal_cols_kernel's downward pass, so the two kernels agree bit for bit.
// Stages with FIR column filters run hal_cols_kernel (their windows would not
// fit beside the two register arrays).
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));  // a fresh value per pass: per-slot addresses and masks are not shared
    return v;
}

__global__ __launch_bounds__(kHalColWidth * kHalColChunks) void hal_cols_reg_kernel(HalArgs a) {
    constexpr int R = kHalColRegRows;
    const int ch = blockIdx.y;
    const HalStageDev* st = a.stage + ch;
    const int count = st->count;
    if (count == 0) return;
    __shared__ float scan[(kHalColChunks * kHalColWidth / 64) * kHalColWidth * 3];
    __shared__ float first_row[kHalColWidth];
    const int cx = threadIdx.x % kHalColWidth;
    const int j = threadIdx.x / kHalColWidth;
    const int W = a.width, H = a.height, L = a.col_len;
    const int x = blockIdx.x * kHalColWidth + cx;
    const bool live = x < W;
    const int xc = live ? x : W - 1;
    const int hi = H - 1 - j * L;
    const int lo = max(0, H - (j + 1) * L);
    const int len = max(0, hi - lo + 1);
    const int m0 = R - len;
    const int top = (H + L - 1) / L - 1;
    auto elem = [&](auto* plane, int r0, int m) {
        const int r = min(max(r0 + m, 0), H - 1);
        return plane + static_cast<size_t>(r) * W + xc;
    };
    const float* Xin = a.in[ch];
    float* Xout = a.out[ch];
    const float c0 = st->c0;
    float acc[R];
    {
        const int r0 = opaque(hi) - (R - 1);
#pragma unroll
        for (int m = 0; m < R; ++m) acc[m] = c0 * *elem(Xin, r0, m);
    }
    for (int k = 0; k < count; ++k) {
        const HalGaussDev* g = st->g + k;
        const float om = st->omega[k];
        const int kind = g->kind;
        if (kind == kHalIdentity) {
            const int r0 = opaque(hi) - (R - 1);
#pragma unroll
            for (int m = 0; m < R; ++m) acc[m] = fmaf(om, *elem(Xin, r0, m), acc[m]);
        } else {
            const Rec rec = load_rec(g);
            const float* tab = &g->col_pow[0][0];
            const float* Q = a.q[ch][k];
            float v[R];
            {
                const int r0 = opaque(hi) - (R - 1);
#pragma unroll
                for (int m = 0; m < R; ++m) v[m] = *elem(Q, r0, m);
            }
            float s[3], e[3], prev[3];
            // Bottom to top, seeded with the bottom row (slot R - 1 of chunk 0).
            const float bottom = v[R - 1];
            if (j == 0) hal_seed(rec, bottom, s);
            else hal_zero(s);
            e[0] = s[0], e[1] = s[1], e[2] = s[2];
            {
                const int mb = opaque(m0);
#pragma unroll
                for (int m = R - 1; m >= 0; --m) {
                    hal_step(rec, v[m], s);
                    const bool at = m == mb;
                    e[0] = at ? s[0] : e[0];
                    e[1] = at ? s[1] : e[1];
                    e[2] = at ? s[2] : e[2];
                }
            }
            chunk_scan<kHalColWidth>(scan, cx, e, tab, prev);
            if (j == 0) hal_seed(rec, bottom, s);
            else s[0] = prev[0], s[1] = prev[1], s[2] = prev[2];
            float last = 0.0f;
            {
                const int mb = opaque(m0);
#pragma unroll
                for (int m = R - 1; m >= 0; --m) {
                    v[m] = hal_step(rec, v[m], s);
                    last = m == mb ? v[m] : last;
                }
            }
            // Row 0 of the upward result: the last output of the top chunk.
            if (j == top) first_row[cx] = last;
            __syncthreads();
            const float first = first_row[cx];
            // Top to bottom, seeded with row 0 of the upward result.
            float init[3];
            if (j == top) hal_seed(rec, first, init);
            else hal_zero(init);
            s[0] = init[0], s[1] = init[1], s[2] = init[2];
            {
                const int mb = opaque(m0);
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const bool at = m == mb;
                    s[0] = at ? init[0] : s[0];
                    s[1] = at ? init[1] : s[1];
                    s[2] = at ? init[2] : s[2];
                    hal_step(rec, v[m], s);
                }
            }
            if (len == 0) hal_zero(s);  // an empty chunk passes its incoming state on unchanged
            chunk_scan_rev<kHalColWidth>(scan, cx, s, tab, prev);
            if (j == top) hal_seed(rec, first, init);
            else init[0] = prev[0], init[1] = prev[1], init[2] = prev[2];
            s[0] = init[0], s[1] = init[1], s[2] = init[2];
            {
                const int mb = opaque(m0);
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const bool at = m == mb;
                    s[0] = at ? init[0] : s[0];
                    s[1] = at ? init[1] : s[1];
                    s[2] = at ? init[2] : s[2];
                    acc[m] = fmaf(om, hal_step(rec, v[m], s), acc[m]);
                }
            }
        }
    }
    if (live) {
        const int r0 = opaque(hi) - (R - 1);
        const int mb = opaque(m0);
#pragma unroll
        for (int m = 0; m < R; ++m)
            if (m >= mb) *elem(Xout, r0, m) = acc[m];
    }
}

// FIR-only stage: grid (ceil(W/64), ceil(H/32), 3), 256 threads. The tile and
// a 9-pixel reflected halo sit in LDS; for every Gaussian the horizontal taps
// fill a 50 x 64 intermediate and the vertical taps accumulate the 64 x 32
// outputs in registers (8 per thread).
constexpr int kTileW = 64;
constexpr int kTileH = 32;
constexpr int kHalo = 9;
constexpr int kInW = kTileW + 2 * kHalo;  // 82
constexpr int kInH = kTileH + 2 * kHalo;  // 50
constexpr int kStrip = 8;                 // outputs per thread and pass

// One FIR Gaussian of radius kR (2 kR + 1 taps, w = g->w) on a tile: the
// horizontal taps over the rows the vertical taps need (kTileH + 2 kR) into
// th, then the vertical taps into acc. Only the real taps are evaluated; the
// sums run in the order of the zero-padded 19-tap form, whose extra terms
// add nothing.
template <int kR>
__device__ __forceinline__ void fir_gaussian(const HalGaussDev* g, float om, const float (*tin)[kInW + 1],
                                             float (*th)[kTileW + 1], int vc, int vr0, float acc[kStrip]) {
    constexpr int kTaps = 2 * kR + 1;
    constexpr int kRows = kTileH + 2 * kR;
    constexpr int kRow0 = kHalo - kR;  // first tin / th row the vertical taps read
    float w[kTaps];
#pragma unroll
    for (int i = 0; i < kTaps; ++i) w[i] = g->w[i];
    for (int item = threadIdx.x; item < kRows * (kTileW / kStrip); item += 256) {
        const int r = kRow0 + item / (kTileW / kStrip);
        const int c = (item % (kTileW / kStrip)) * kStrip;
        float win[kStrip + kTaps - 1];
#pragma unroll
        for (int i = 0; i < kStrip + kTaps - 1; ++i) win[i] = tin[r][c + kHalo - kR + i];
#pragma unroll
        for (int i = 0; i < kStrip; ++i) {
            float sum = 0.0f;
#pragma unroll
            for (int o = 0; o < kTaps; ++o) sum = fmaf(win[i + o], w[o], sum);
            th[r][c + i] = sum;
        }
    }
    __syncthreads();
    float win[kStrip + kTaps - 1];
#pragma unroll
    for (int i = 0; i < kStrip + kTaps - 1; ++i) win[i] = th[kRow0 + vr0 + i][vc];
#pragma unroll
    for (int i = 0; i < kStrip; ++i) {
        float sum = 0.0f;
#pragma unroll
        for (int o = 0; o < kTaps; ++o) sum = fmaf(win[i + o], w[o], sum);
        acc[i] = fmaf(om, sum, acc[i]);
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void hal_fir_tile_kernel(HalArgs a) {
    const int ch = blockIdx.z;
    const HalStageDev* st = a.stage + ch;
    const int count = st->count;
    const int W = a.width, H = a.height;
    const int x0 = blockIdx.x * kTileW, y0 = blockIdx.y * kTileH;
    const float* X = a.in[ch];
    float* O = a.out[ch];
    __shared__ float tin[kInH][kInW + 1];
    __shared__ float th[kInH][kTileW + 1];
    const int t = threadIdx.x;
    batched<8, float>(
        t, kInH * kInW, 256,
        [&](int i) {
            const int r = i / kInW, c = i - r * kInW;
            return X[static_cast<size_t>(hal_reflect(y0 - kHalo + r, H)) * W + hal_reflect(x0 - kHalo + c, W)];
        },
        [&](int i, float v) {
            const int r = i / kInW, c = i - r * kInW;
            tin[r][c] = v;
        });
    __syncthreads();
    // Vertical-pass ownership: column vc, rows vr0 .. vr0+7.
    const int vc = t % kTileW;
    const int vr0 = (t / kTileW) * kStrip;
    float acc[kStrip];
    const float c0 = count ? st->c0 : 1.0f;
#pragma unroll
    for (int i = 0; i < kStrip; ++i) acc[i] = c0 * tin[kHalo + vr0 + i][kHalo + vc];
    for (int k = 0; k < count; ++k) {
        const HalGaussDev* g = st->g + k;
        const float om = st->omega[k];
        if (g->kind == kHalIdentity) {
#pragma unroll
            for (int i = 0; i < kStrip; ++i) acc[i] = fmaf(om, tin[kHalo + vr0 + i][kHalo + vc], acc[i]);
            continue;
        }
        switch (g->radius) {
            case 0: fir_gaussian<0>(g, om, tin, th, vc, vr0, acc); break;
            case 1: fir_gaussian<1>(g, om, tin, th, vc, vr0, acc); break;
            case 2: fir_gaussian<2>(g, om, tin, th, vc, vr0, acc); break;
            case 3: fir_gaussian<3>(g, om, tin, th, vc, vr0, acc); break;
            case 4: fir_gaussian<4>(g, om, tin, th, vc, vr0, acc); break;
            case 5: fir_gaussian<5>(g, om, tin, th, vc, vr0, acc); break;
            case 6: fir_gaussian<6>(g, om, tin, th, vc, vr0, acc); break;
            case 7: fir_gaussian<7>(g, om, tin, th, vc, vr0, acc); break;
            case 8: fir_gaussian<8>(g, om, tin, th, vc, vr0, acc); break;
            default: fir_gaussian<9>(g, om, tin, th, vc, vr0, acc); break;
        }
    }
    const int x = x0 + vc;
    if (x < W) {
#pragma unroll
        for (int i = 0; i < kStrip; ++i) {
            const int y = y0 + vr0 + i;
            if (y < H) O[static_cast<size_t>(y) * W + x] = acc[i];
        }
    }
}

size_t halation_row_lds_bytes(int width) { return static_cast<size_t>(2) * width * sizeof(float); }

hipError_t halation_kernels_init() {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(hal_rows_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(halation_row_lds_bytes(kHalMaxLine)));
}

hipError_t launch_halation_stage(const HalArgs& a, hipStream_t stream) {
    if (a.width <= 0 || a.height <= 0 || a.width > kHalMaxLine) return hipErrorInvalidValue;
    if (a.row_len * kHalRowThreads < a.width || a.col_len * kHalColChunks < a.height) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hal_rows_kernel, dim3(a.height, 3), dim3(kHalRowThreads), halation_row_lds_bytes(a.width),
                       stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid((a.width + kHalColWidth - 1) / kHalColWidth, 3);
    if (a.col_len <= kHalColRegRows && !a.column_fir)
        hipLaunchKernelGGL(hal_cols_reg_kernel, grid, dim3(kHalColWidth * kHalColChunks), 0, stream, a);
    else
        hipLaunchKernelGGL(hal_cols_kernel, grid, dim3(kHalColWidth * kHalColChunks), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_halation_fir_stage(const HalArgs& a, hipStream_t stream) {
    if (a.width <= 0 || a.height <= 0) return hipErrorInvalidValue;
    for (int c = 0; c < 3; ++c)
        if (a.in[c] == a.out[c]) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hal_fir_tile_kernel, dim3((a.width + kTileW - 1) / kTileW, (a.height + kTileH - 1) / kTileH, 3),
                       dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace fjx
