// accumulate.h -- launch interface of the accumulate kernels.
#pragma once

#include "planner.h"

namespace ionp {

// in (already in the output dtype) -> out along `axis` (normalised)
ionp_status run_accumulate(int op, const ionp_array_desc *in, int axis, const ionp_array_desc *out,
                           const ionp_config *cfg, hipStream_t s);

}  // namespace ionp
