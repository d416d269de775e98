// Persistent full-matrix kernels, without a query bias, with end cells: one per register class.
#include "full_kernel.h"

namespace saber {

hipError_t launch_full_multi_nobias_ends(int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s) {
    if (nt == 768) return launch_full_class<false, true, 768>(a, grid, s);
    if (nt == 512) return launch_full_class<false, true, 512>(a, grid, s);
    return hipErrorInvalidValue;
}

int full_multi_blocks_nobias_ends(int nt) {
    if (nt == 768) return full_class_blocks_per_cu<false, true, 768>();
    if (nt == 512) return full_class_blocks_per_cu<false, true, 512>();
    return 1;
}

}  // namespace saber
