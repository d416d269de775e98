// Compiled variants of the packed band kernel, part d of the table
// (L lanes per job pair, C band cells per lane, K capped trailing cells; the
// band width W = L*C serves clipped widths W-K+1 .. W).
#include "band_kernel.h"

namespace saber {

#define SABER_VARIANT(L, C, K) \
    {L, C, K, 0, &launch_band<L, C, K, false>}, {L, C, K, 1, &launch_band<L, C, K, true>},

static const BandVariant kPart[] = {
SABER_VARIANT(4, 57, 16)
SABER_VARIANT(4, 61, 16)
SABER_VARIANT(4, 65, 16)
SABER_VARIANT(4, 69, 16)
SABER_VARIANT(4, 73, 16)
};

const BandVariant* band_variants_part_d(int* count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace saber
