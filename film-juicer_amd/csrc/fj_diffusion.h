// fj_diffusion.h — camera/enlarger diffusion (device blocks).
//
// A diffusion stage blends each plane with its convolution by a radial PSF of
// nine exponentials under the whole-sample mirror boundary
// (JuicerCudaDiffusion.cu:403-494):
//   out = (1 - f) x + f (psf * reflect101(x)).
// The PSF support is (2r + 1)^2 with r up to half the frame, so the
// convolution runs in the DCT-II domain of a padded plane: the plane is
// extended by r samples of its reflect-101 mirror on every side (padded to a
// 2^a 3^b 5^c length per axis), and the DCT-II basis of that padded plane
// diagonalizes convolution by the (even, even) PSF with eigenvalues
//   T(u, v) = sum_{dx, dy in [-r, r]} psf(dx, dy) cos(pi u dx / PW) cos(pi v dy / PH).
// Every output sample reads inputs at most r away, all inside the padded
// plane, so the padded plane's own boundary never reaches the frame's
// samples. T is built once per PSF and padded size by two FP64 cosine
// products: M[dy][u] = sum_dx q[dy][dx] C_W[dx][u], then
// T[u][v] = sum_dy M[dy][u] C_H[dy][v], with q the PSF quadrant and
// C_N[d][j] = c_d cos(pi j d / N), c_0 = 1, c_d = 2.
#pragma once

#include <hip/hip_runtime.h>

namespace fjx {

// C[d][j] = c_d cos(pi (d j mod 2n) / n) for d <= rows - 1, j < n, from the
// FP64 table cos_table[m] = cos(pi m / n), m < 2n.
struct DiffCosArgs {
    const double* cos_table;
    double* out;
    int rows;
    int n;
};

// out[c][a][b] = sum_k L[c][k][a] * R[k][b] for a < M, b < N, k < K, per
// channel c (blockIdx.z). L has leading dimension M, R and out N.
template <typename TOut>
struct AtbArgs {
    const double* L[3];
    const double* R;
    TOut* out[3];
    int M, N, K;
};

hipError_t launch_diffusion_cos(const DiffCosArgs& a, hipStream_t s);
hipError_t launch_diffusion_atb(const AtbArgs<double>& a, int channels, hipStream_t s);
hipError_t launch_diffusion_atb(const AtbArgs<float>& a, int channels, hipStream_t s);

}  // namespace fjx
