// batch.hip -- placeholder
#include "planner.h"
using namespace ionp;
extern "C" {
ionp_status ionp_batch_create(const ionp_call *, int32_t, void *, ionp_batch **) {
    return set_error(IONP_ERR_UNSUPPORTED, "batching not built yet");
}
ionp_status ionp_batch_run(ionp_batch *, int32_t, void *) { return set_error(IONP_ERR_UNSUPPORTED, "batching not built yet"); }
ionp_status ionp_batch_info(const ionp_batch *, int64_t *, int64_t *) { return set_error(IONP_ERR_UNSUPPORTED, "batching not built yet"); }
void ionp_batch_destroy(ionp_batch *) {}
}
