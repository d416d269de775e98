// Compiled variants of the packed band kernel, part a of the table
// (L lanes per job pair, C band cells per lane, K capped trailing cells; the
// band width W = L*C serves clipped widths W-K+1 .. W).
#include "band_kernel.h"

namespace saber {

#define SABER_VARIANT(L, C, K) \
    {L, C, K, 0, &launch_band<L, C, K, false>}, {L, C, K, 1, &launch_band<L, C, K, true>},

static const BandVariant kPart[] = {
SABER_VARIANT(1, 9, 8)
SABER_VARIANT(1, 17, 8)
SABER_VARIANT(1, 25, 8)
SABER_VARIANT(1, 33, 8)
SABER_VARIANT(1, 41, 8)
SABER_VARIANT(1, 49, 8)
SABER_VARIANT(1, 57, 8)
SABER_VARIANT(1, 65, 8)
};

const BandVariant* band_variants_part_a(int* count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace saber
