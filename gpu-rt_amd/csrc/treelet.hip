// treelet.hip -- SAH treelet restructuring of the binary LBVH (Karras & Aila, HPG 2013).
#include "gpurt_internal.h"

namespace gpurt_impl {

int run_treelets(BuildBufs&, int, int, hipStream_t) {
    return set_error(GPURT_E_UNSUPPORTED, "treelet refinement is not available in this build");
}

}  // namespace gpurt_impl
