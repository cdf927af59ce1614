// advisor.h -- the learned launch configuration consulted by the entry
// points for calls made without one (advisor.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "ionp_common.h"

namespace ionp {

// A call the advisor times for a trial: begin() records the start event
// right before the call's launches, the destructor the end event after them
// (a call that returns before begin() records neither and gives no sample).
struct AdviseTiming {
    void *trial = nullptr;
    uint64_t generation = 0;
    int arm = 0, slot = 0;
    hipEvent_t start = nullptr, end = nullptr;
    hipStream_t stream = nullptr;
    bool started = false;
    void begin();
    // the call failed after begin(): its slot is freed without a sample
    ionp_status done(ionp_status s) {
        if (s != IONP_OK) started = false;
        return s;
    }
    ~AdviseTiming();
};

// The cost model's configuration for this call's layout, written to *slot
// and returned; nullptr (the built-in heuristic) when no model is loaded,
// the layout's proposal is not ready yet (it is then queued for inference)
// or its trial times the heuristic on this call.  *timing is armed when the
// call is one of its layout's timed trial calls.
const ionp_config *advise(int kind, int op, int nin, const ionp_array_desc *const *ins, const ionp_array_desc *out,
                          const int32_t *axes, int32_t naxes, int32_t keepdims, hipStream_t s, ionp_config *slot,
                          AdviseTiming *timing);

}  // namespace ionp
