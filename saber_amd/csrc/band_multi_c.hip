// The multi-shape packed band kernel without query bias, scores only.
#include "band_kernel.h"

namespace saber {

hipError_t launch_band_multi_nobias_scores(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    SABER_LAUNCH((band_multi_kernel<false, false>), dim3(nblocks), dim3(768), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
