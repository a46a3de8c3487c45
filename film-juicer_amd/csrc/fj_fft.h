// fj_fft.h — device argument blocks of the LDS Stockham FFT / DCT kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace fjx {

constexpr int kMaxFftPasses = 24;
constexpr int kMaxDctComponents = 8;
constexpr int kMaxLdsFft = 16384;      // in-place engine: 16384 float2 = 128 KiB of the 160 KiB LDS
constexpr int kMaxPingPongFft = 8192;  // ping-pong engine: 2 x 8192 float2
constexpr int kMaxRowThreads = 1024;

struct FftPlanDev {
    int n = 0;
    int passes = 0;
    int radix[kMaxFftPasses] = {};
    // Per pass p (ns = product of the radices before p): its twiddles
    // w_n^(r k n/(ns R)) for k < ns, 1 <= r < R sit at table[n + tw_offset[p] +
    // k (R-1) + r - 1], so neighbouring butterflies read neighbouring words;
    // ns_magic[p] = ceil(2^32 / ns) divides butterfly indices by ns exactly.
    int tw_offset[kMaxFftPasses] = {};
    unsigned ns_magic[kMaxFftPasses] = {};
    // In-place engine: pass p's base twiddles w_n^(k n/(ns R)) for k < ns at
    // table[n + tw1_offset[p] + k]; a butterfly loads one and forms the other
    // powers by squaring and multiplying (at most 2 log2 R products).
    int tw1_offset[kMaxFftPasses] = {};
    // In-place engine (2,3,5-smooth n): radices from {2,3,4,5,6,8,9,10,12,15,
    // 16,18,20,24,25}, one butterfly per thread, `threads` >= n / radix of every pass.
    int inplace = 0;
    int threads = 0;
    // Bluestein plan (n not 2,3,5-smooth): the length-n DFT is the chirp-z
    // convolution X_k = c_k sum_j (x_j c_j) b_(k-j), c_j = exp(-i pi j^2 / n),
    // b = conj(c), evaluated with two in-place transforms of the smooth
    // length m >= 2n - 1 that the pass fields above describe (their twiddles
    // relative to table + m). table[chirp_offset + j] = c_j for j < n and
    // table[spectrum_offset + k] = the length-m DFT of the wrapped b, k < m.
    int bluestein = 0;
    int m = 0;
    int chirp_offset = 0;
    int spectrum_offset = 0;
};

// Length of the transform the pass fields run (m for a Bluestein plan).
__host__ __device__ inline int fft_length(const FftPlanDev& plan) { return plan.bluestein ? plan.m : plan.n; }

// Up to three planes transformed by one launch (blockIdx.y selects the plane).
struct DctRowsArgs {
    FftPlanDev plan;
    const float2* table = nullptr;    // exp(-2 pi i e / n), e < n
    const float2* half_tw = nullptr;  // exp(-i pi k / (2n)), k < n
    const float* in[3] = {};
    float* out[3] = {};
    int rows = 0;           // rows [row_begin, rows) are transformed
    int row_begin = 0;
    long long in_stride = 0;
    long long out_stride = 0;
    // Padded-line mode (pad >= 0) for rows of src_w samples:
    //  forward rows read the reflect-101 extension of row y, element i
    //    <- in[y * in_stride + reflect101(i - pad, src_w)];
    //  inverse rows keep only the row's own samples and blend them with it:
    //    out[y * out_stride + i - pad] = (1 - fraction) * mix[same] + fraction * value.
    int pad = -1;
    int src_w = 0;
    const float* mix[3] = {};
    float fraction = 0.0f;
    // Filter rows: only elements [out_begin, out_begin + out_count) of each
    // transformed row are written, at out[row * out_stride + i - out_begin]
    // (out_count 0: the whole row at offset 0).
    int out_begin = 0;
    int out_count = 0;
};

// Separable-sum transfer function T(row, k) = sum_c weight[c] * t_rows[c][row] * t_cols[c][k],
// or, when full[] is set (for every plane of the launch), the table
// T(row, k) = full[plane][row * n + k].
struct DctTransfer {
    int count = 0;
    float weight[kMaxDctComponents] = {};
    const float* t_rows[kMaxDctComponents] = {};
    const float* t_cols[kMaxDctComponents] = {};
    const float* full[3] = {};
};

struct TransposeArgs {
    const float* in[3] = {};
    float* out[3] = {};
    int width = 0;   // of the input
    int height = 0;  // of the input
    long long in_stride = 0;
    long long out_stride = 0;
    // src_rows > 0: input row y is read from row reflect101(y - row_pad,
    // src_rows) (the reflect-101 row extension of a src_rows-row plane).
    int src_rows = 0;
    int row_pad = 0;
};

// Host-side plan. A 2,3,5-smooth n gets the in-place engine's composite
// radices (fewest passes, then the largest smallest radix, so the fewest
// threads), odd radices first; any other n a Bluestein plan over the
// smallest smooth m >= 2n - 1 (n up to kMaxLdsFft / 2).
bool make_fft_plan(int n, FftPlanDev& plan);
// Smallest length >= n that has an in-place plan (0 if none up to kMaxLdsFft).
int inplace_fft_length(int n);
// exp(-2 pi i e / L) for e < L = fft_length(plan of n), then the per-pass
// twiddles of the plan, in float2 pairs (then a Bluestein plan's chirp and
// spectrum); and exp(-i pi k / 2n) for k < n.
std::vector<float> fft_twiddles(int n);
std::vector<float> dct_half_twiddles(int n);

size_t dct_lds_bytes(const FftPlanDev& plan);
hipError_t launch_dct2_rows(const DctRowsArgs& a, int planes, hipStream_t s);
hipError_t launch_dct3_rows(const DctRowsArgs& a, int planes, hipStream_t s);
hipError_t launch_dct_filter_rows(const DctRowsArgs& a, const DctTransfer& tf, int planes, hipStream_t s);
hipError_t launch_transpose(const TransposeArgs& t, int planes, hipStream_t s);
hipError_t fft_kernels_init(int* max_lds_n);

}  // namespace fjx
