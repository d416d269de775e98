// Variant table of the packed band kernel (every compiled shape, without and
// with bias) and the launcher of the multi-shape kernel.
#include <vector>

#include "band_kernel.h"

namespace saber {

hipError_t launch_band_multi_nobias_ends(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s);
hipError_t launch_band_multi_bias_ends(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s);
hipError_t launch_band_multi_nobias_scores(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s);
hipError_t launch_band_multi_bias_scores(const BandKernelArgs& a, uint32_t nblocks, hipStream_t s);

const BandVariant* band_variants(int* count) {
    static const std::vector<BandVariant> table = [] {
        std::vector<BandVariant> t;
#define SABER_BAND_ENTRY(L_, C_, K_) t.push_back({L_, C_, K_, 0}); t.push_back({L_, C_, K_, 1});
        SABER_BAND_SHAPES(SABER_BAND_ENTRY)
#undef SABER_BAND_ENTRY
        return t;
    }();
    *count = (int)table.size();
    return table.data();
}

hipError_t launch_band_multi(const BandKernelArgs& a, uint32_t nblocks, bool bias, bool ends, hipStream_t s) {
    if (ends) return bias ? launch_band_multi_bias_ends(a, nblocks, s) : launch_band_multi_nobias_ends(a, nblocks, s);
    return bias ? launch_band_multi_bias_scores(a, nblocks, s) : launch_band_multi_nobias_scores(a, nblocks, s);
}

}  // namespace saber
