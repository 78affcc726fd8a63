// refit.hip -- BVH refit for moved vertices (topology unchanged).
#include "gpurt_internal.h"

namespace gpurt_impl {

int launch_refit(gpurt_bvh_s*, const float*, size_t, hipStream_t) {
    return set_error(GPURT_E_UNSUPPORTED, "refit is not available in this build");
}

}  // namespace gpurt_impl
