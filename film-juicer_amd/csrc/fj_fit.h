// fj_fit.h — the differentiable print route: launch arguments shared by the
// forward and adjoint kernels (fj_fit.hip) and their host (fj_fit.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "fj_device.h"

namespace fjx {

// Parameter slots of the gradient (FjFitParams order).
enum FitSlot : int { kFitEv = 0, kFitPrintExposure, kFitFilterY, kFitFilterM, kFitFilmGamma, kFitPrintGamma, kFitDirK,
                     kFitUsed = 7, kFitSlots = 8 };

struct FitArgs {
    PixelParams pp;          // the parameters' record, assembled as a render assembles it
    const float* src;        // packed RGBA source rows
    float* dst;              // packed RGBA destination rows (forward)
    const float* delta;      // grain increment C/M/Y planes, or null
    float* saved;            // [log raw B/G/R][filtered DIR corrections][densities C/M/Y], 3 planes each
    const float* grad_dst;   // packed RGBA loss gradient (backward)
    float* grad_src;         // packed RGBA source gradient (backward)
    float* g_dens;           // 3 planes: gradient of the grained densities C/M/Y
    float* g_filtered;       // 3 planes: gradient of the filtered corrections (then of the raw ones)
    float* g_lograw;         // 3 planes: gradient of the log raw exposure B/G/R
    float* partials;         // [launch][block][kFitSlots] parameter partial sums
    int partial_blocks;      // blocks per adjoint launch (the partials' stride)
    // d ln(filtered illuminant) / d cc for M and Y per compacted print
    // wavelength (the order of pp.pr.spec): {aM, aY, 0, 0}.
    const float4* alpha;
    // d(compensated axis) / d dir_k, packed like the film curves' axes:
    // [B | G | R], n floats each.
    const float* daxis;
    float mallett_A[9];      // Mallett exposure E_b = sum_j A[b][j] s_j (b in B, G, R) before the mid-gray scale
    float donor_k_unit[3];   // d(donor K) / d dir_k, B/G/R
    float print_gamma;
    float print_route_scale; // d(outer scale) / d(print exposure)
    float film_gamma;
};

// Blocks of every fit launch over `total` pixels (the partials' row count
// per adjoint launch).
int fit_blocks(long long total);
hipError_t launch_fit_print(const FitArgs& a, hipStream_t stream);
// Adjoint stage 0 (print), 1 (DIR final) or 2 (source).
hipError_t launch_fit_adjoint(const FitArgs& a, int stage, hipStream_t stream);
// Sums `rows` partial rows into out[kFitSlots].
hipError_t launch_fit_reduce(const float* partials, int rows, float* out, hipStream_t stream);

}  // namespace fjx
