// Compiled variants of the packed band kernel, part c of the table
// (L lanes per job pair, C band cells per lane, K capped trailing cells; the
// band width W = L*C serves clipped widths W-K+1 .. W).
#include "band_kernel.h"

namespace saber {

#define SABER_VARIANT(L, C, K) \
    {L, C, K, 0, &launch_band<L, C, K, false>}, {L, C, K, 1, &launch_band<L, C, K, true>},

static const BandVariant kPart[] = {
SABER_VARIANT(4, 37, 16)
SABER_VARIANT(4, 41, 16)
SABER_VARIANT(4, 45, 16)
SABER_VARIANT(4, 49, 16)
SABER_VARIANT(4, 53, 16)
};

const BandVariant* band_variants_part_c(int* count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace saber
