// treelet.hip -- SAH treelet restructuring of the binary LBVH (Karras & Aila, HPG 2013).
#include "gpurt_internal.h"

namespace gpurt_impl {

int run_treelets(int, int, int, const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, uint2*,
                 uint32_t*, uint32_t*, const uint8_t*, const float4*, float4*, uint32_t*,
                 hipStream_t) {
    return set_error(GPURT_E_UNSUPPORTED, "treelet refinement is not available in this build");
}

}  // namespace gpurt_impl
