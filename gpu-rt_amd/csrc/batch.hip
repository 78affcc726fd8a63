// batch.hip -- batched builds and queries over many meshes.
#include "gpurt_internal.h"

using namespace gpurt_impl;

extern "C" {

int gpurt_batch_create(gpurt_ctx, const gpurt_batch_mesh*, size_t, const gpurt_build_opts*,
                       const gpurt_query_opts*, gpurt_batch* out) {
    if (out) *out = nullptr;
    return set_error(GPURT_E_UNSUPPORTED, "batched path is not available in this build");
}
int gpurt_batch_run(gpurt_batch, void*) {
    return set_error(GPURT_E_UNSUPPORTED, "batched path is not available in this build");
}
void gpurt_batch_destroy(gpurt_batch) {}

}  // extern "C"
