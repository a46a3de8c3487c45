// fj_tclut.h — the film TC-LUT build on the GPU (device launches).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fjx {

constexpr int kTcSide = 192;
constexpr int kTcCells = kTcSide * kTcSide;
constexpr int kTcSamples = 81;

// Failure bits of the build's device error words (the first failing cell in
// the reference's loop order is kept per stage through atomicMin of
// cell * 8 + rank; rank orders a cell's checks).
constexpr unsigned int kTcNoFailure = 0xFFFFFFFFu;

struct TcBlurArgs {
    const float* src;     // 192 x 192 x 81
    float* dst;
    const float* kernel;  // 2 radius + 1 weights
    int radius;
    unsigned int* first_failure;
};

struct TcIntegrateArgs {
    const float* spectra;        // 192 x 192 x 81 (blurred or not)
    const float* sensitivity;    // 81 x 3
    const float* illuminant;     // 81, multiplied into the energy when relight != 0
    int relight;
    float denominator[3];
    const float* surface_factor; // 192 x 192 x 3 exp2 factors, or null
    float* out;                  // 192 x 192 x 4
    unsigned int* first_failure;
};

struct TcRemapArgs {
    const float* integrated;     // 192 x 192 x 4
    const float* hull_xy;        // 1025 x 2
    float hull_center[2];
    const float* roots;          // 192: sqrt(c / 191), from the host
    const float* steps;          // 192: m / 191
    float* out;                  // 192 x 192 x 4
    unsigned int* first_failure;
};

hipError_t launch_tc_blur(const TcBlurArgs& a, hipStream_t s);
hipError_t launch_tc_integrate(const TcIntegrateArgs& a, hipStream_t s);
hipError_t launch_tc_remap(const TcRemapArgs& a, hipStream_t s);

}  // namespace fjx
