// fj_host_imports.h — optional host-defined functions, resolved at run time.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_diffusion_host.h"

namespace fjx {

bool host_defines_diffusion_components();
bool host_defines_dir_boundary();

// The nine PSF components of a diffusion stage: the host's
// fj_diffusion_components_prepare when the process defines it, else the
// native expansion. On failure `category` and `diagnostic` are the frame's.
bool stage_components(const FjDiffusion& d, double lambda_px[kDiffusionComponents],
                      double weights[3][kDiffusionComponents], uint32_t& category, std::string& diagnostic);

// When the process defines the DIR boundary functions: every recursive
// spatial-DIR component's reflected boundary on a width x height frame must
// be certified by the host (describe, then acquire/view/release for
// identities not certified before). True when the host defines none.
bool host_dir_boundaries(const FjSpatialDir& sd, int width, int height, uint32_t& category, std::string& diagnostic);

}  // namespace fjx
