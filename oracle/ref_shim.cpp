// ref_shim.cpp — C entry points over reference sources compiled where they
// lie under /root/reference (test infrastructure: the checker of the oracle,
// built into oracle/_ref/ only in a container that holds the reference).
//
//   tests/scanner/output_encoding_oracle.cpp   ScannerOracle::apply_encoding
//   src/GamutCompression.h                     Gamut::compress_output_oklch
//   third_party/openrand/philox.h              openrand::Philox (the glare
//                                              and grain fallback streams)
#include <cstdint>

#include <openrand/philox.h>

#include "GamutCompression.h"
#include "OutputColor.h"
#include "output_encoding_oracle.h"

extern "C" {

// Output matrix + CCTF + clamp of one RGB triple for catalogue entry
// `color_space` (OutputEncoding::ColorSpace index).
void fjr_apply_encoding(int color_space, int apply_cctf, int input_is_output_space, double rgb[3]) {
    OutputEncoding::Params p;
    p.colorSpace = OutputEncoding::colorSpaceFromIndex(color_space);
    p.applyCctfEncoding = apply_cctf != 0;
    p.inputIsOutputSpace = input_is_output_space != 0;
    ScannerOracle::apply_encoding(p, rgb);
}

// OkLch output compression of one RGB triple (FP32, the reference's shared
// host/device function).
void fjr_compress_output_oklch(const float* cmax, const float entry[9], const float exit[9],
                               const float lightness_knee[3], const float chroma_knee[3], float rgb[3]) {
    Gamut::OutputCompressionView v;
    v.cmax = cmax;
    v.nativeRgbToD65Xyz = entry;
    v.d65XyzToNativeRgb = exit;
    v.oklabXyzToLms = Gamut::kOklabXyzToLms.data();
    v.oklabLmsToXyz = Gamut::kOklabLmsToXyz.data();
    v.oklabLmsRootToLab = Gamut::kOklabLmsRootToLab.data();
    v.oklabLabToLmsRoot = Gamut::kOklabLabToLmsRoot.data();
    v.lightnessKnee = {lightness_knee[0], lightness_knee[1], lightness_knee[2]};
    v.chromaKnee = {chroma_knee[0], chroma_knee[1], chroma_knee[2]};
    Gamut::compress_output_oklch(v, rgb);
}

// `count` successive rand<uint32_t>() draws of openrand::Philox(seed, ctr,
// global_seed, ctr1), and `count` rand<float>() draws of a second generator
// with the same arguments.
void fjr_philox_draws(uint64_t seed, uint32_t ctr, uint32_t global_seed, uint32_t ctr1, int count, uint32_t* words,
                      float* floats) {
    openrand::Philox a(seed, ctr, global_seed, ctr1);
    for (int i = 0; i < count; ++i) words[i] = a.rand<uint32_t>();
    openrand::Philox b(seed, ctr, global_seed, ctr1);
    for (int i = 0; i < count; ++i) floats[i] = b.rand<float>();
}

}  // extern "C"
