// elementwise.h -- launch interface of the elementwise kernels.
#pragma once

#include "planner.h"

namespace ionp {

// rows x inner fast-path arguments (dim 0 of the coalesced space is `inner`)
struct RowArgs {
    int64_t inner, rows, tiles_per_row;
    FastDiv tpr;
    int32_t nouter;
    int64_t oshape[kMaxKDims];
    FastDiv odiv[kMaxKDims];
    int64_t ost[2][kMaxKDims], ost_out[kMaxKDims];
    int64_t inner_st[2];
    int32_t pair_ok[2];
    const char *in[2];
    char *out;
    uint32_t *fpe;  // IEEE flag word (FPE variants only)
    int32_t nt_loads;  // stream the inputs with nontemporal loads
};

int num_cus();
// MI355X's Infinity Cache (last level, shared by all XCDs): inputs larger
// than this are streamed with nontemporal loads
constexpr int64_t kInfinityCacheBytes = 256ll << 20;
ionp_status run_elementwise(EwArgs &A, const ionp_config *cfg, hipStream_t s);
// where / cast over a rows layout (rows_select_cast.hip); false: not eligible
bool try_rows_select_cast(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st);
// any elementwise call with an input contiguous along a dim other than the
// output's contiguous one: that input staged through LDS (tiled_any.hip)
bool try_tiled_any(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st);
ionp_status launch_generic(EwArgs A, const ionp_config *cfg, hipStream_t s);
// IEEE flag tracking of the calling thread: sets it on/off, returns the old state
int fpe_thread_swap(int on);
// FpeOff: the plain kernels for the lifetime of the guard (batches, autotuning)
struct FpeOff {
    int old;
    FpeOff() : old(fpe_thread_swap(0)) {}
    ~FpeOff() { fpe_thread_swap(old); }
};

}  // namespace ionp
