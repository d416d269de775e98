// The multi-shape packed band kernel without query bias, scores and end cells.
#include "band_kernel.h"

namespace saber {

hipError_t launch_band_multi_nobias_ends(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    SABER_LAUNCH((band_multi_kernel<false, true>), dim3(nblocks), dim3(768), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
