// fj_halation.h — scatter/halation on the camera carrier (device blocks).
//
// Each channel runs two linear stages in place on its carrier plane X
// (launch_scatter_halation, JuicerCudaScatterHalation.cu:690-852):
//   scatter: X <- (1-a) X + a [(1-w_t) C X + w_t sum_i A_i T_i X]
//   bounce:  X <- (X + s sum_j b_j G_j X) / (1 + s)   (renormalized)
// Both have the form X <- c0 X + sum_k omega_k G_k X, where every G_k is a
// separable Gaussian: identity, a reflected FIR (<= 19 taps) or a
// replicate-seeded third-order recursion (direct or factored form).
//
// A stage is two launches. The row kernel keeps one row in LDS and writes the
// row-filtered plane Q_k of every non-identity Gaussian. The column kernel
// runs the column filters of all Gaussians and the weighted combination, and
// writes X. Recursions run chunk-parallel: every thread owns one chunk of a
// line, first runs its chunk from rest to get the chunk's end state, then a
// scan over chunks with the precomputed chunk transition powers M^n = A^(L n)
// gives each chunk its true incoming state, and the chunk is re-run from that
// state. This is exact in exact arithmetic. The scan runs inside each wave
// with lane shuffles and across the waves of a block through LDS, so a scan
// costs two barriers.
//
// Every recursion runs in the factored form of JuicerCudaScatterHalation.cu
// :183-198 (real pole p, second-order section y with increment d) in FP32.
// Direct-form descriptors (B, B1, B2, B3) are converted on the host, in FP64,
// to the factored form with the same poles, an output gain kappa (the direct
// form's DC gain, which f32 coefficient rounding moves off 1) and the state
// that reproduces the direct form's "all past outputs = first sample" seed.
// The direct form itself has a rounding-noise gain of 1/B (B ~ 1e-4 at a
// 30 px sigma), so evaluating it in FP32 would drift by ~1e-3 of the signal;
// the factored form stays within a few 1e-7.
#pragma once

#include <hip/hip_runtime.h>

namespace fjx {

constexpr int kHalMaxGauss = 5;
constexpr int kHalRowThreads = 256;  // chunks per row
constexpr int kHalColWidth = 16;     // columns per column-kernel block
constexpr int kHalColChunks = 64;    // chunks per column
constexpr int kHalColRegRows = 36;   // chunk rows a column thread holds in registers (H <= 2304)
constexpr int kHalMaxLine = 8192;    // longest row the row kernel holds in LDS

enum HalKind : int { kHalIdentity = 0, kHalFir = 1, kHalRecursive = 2 };

struct HalGaussDev {
    int kind;
    int radius;
    float w[19];
    float h, g, t;         // factored recursion
    float kappa;           // output gain
    float seed[3];         // state seeded from a pass's first sample: seed * first
    // Chunk transition powers of the state (p, y, d), row-major 3x3:
    // row_pow[n] = (A^Lr)^n for n <= 64 (chunks per wave in the row pass);
    // col_pow[n] = (A^Lc)^n, n <= 4.
    float row_pow[65][9];
    float col_pow[5][9];
};

struct HalStageDev {
    int count;             // Gaussians of this channel's stage (0: plane unchanged)
    float c0;              // weight of the stage input
    float omega[kHalMaxGauss];
    HalGaussDev g[kHalMaxGauss];
};

struct HalArgs {
    const float* in[3];                // stage input planes (R, G, B descriptor order)
    float* out[3];                     // stage output planes (may equal `in` for the row/column pair)
    float* q[3][kHalMaxGauss];         // row-filtered scratch planes
    const HalStageDev* stage;          // [3]
    int width, height;
    int row_len;                       // Lr = ceil(width / kHalRowThreads)
    int col_len;                       // Lc = ceil(height / kHalColChunks)
    int column_fir;                    // some Gaussian of the stage is an FIR
};

// Row + column launches; any Gaussian kinds; `out` may alias `in`.
hipError_t launch_halation_stage(const HalArgs& a, hipStream_t stream);
// One launch for stages whose Gaussians are all FIR or identity; `out` must
// not alias `in` (tiles read a halo of their neighbours).
hipError_t launch_halation_fir_stage(const HalArgs& a, hipStream_t stream);
size_t halation_row_lds_bytes(int width);

}  // namespace fjx
