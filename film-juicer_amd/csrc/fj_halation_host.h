// fj_halation_host.h — host side of the scatter/halation stages.
#pragma once

#include <string>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_halation.h"

namespace fjx {

struct HalationStages {
    HalStageDev scatter[3];
    HalStageDev bounce[3];
    bool scatter_active;
    bool bounce_active;
    int row_len, col_len;
};

// Returns false and names the malformed field.
bool halation_validate(const FjScatterHalation& h, std::string& field);
// Stage tables for a width x height carrier; false when a recursion has no
// stable factored form.
bool halation_stages(const FjScatterHalation& h, int width, int height, HalationStages& out);
hipError_t halation_kernels_init();

}  // namespace fjx
