// Compiled variants of the full-matrix kernel, part e: without a query bias, score only
// (L lanes per subject pair, C subject columns per lane; subjects of length
// up to L*C).
#include "full_kernel.h"

namespace saber {

#define SABER_FULL(L, C) {L, C, 0, 0, &launch_full<L, C, false, false>},

static const FullVariant kPart[] = {
SABER_FULL(1, 16)
SABER_FULL(1, 24)
SABER_FULL(1, 32)
SABER_FULL(1, 40)
SABER_FULL(1, 48)
SABER_FULL(2, 32)
SABER_FULL(2, 40)
SABER_FULL(2, 48)
SABER_FULL(4, 32)
SABER_FULL(4, 40)
SABER_FULL(4, 48)
SABER_FULL(8, 32)
SABER_FULL(8, 40)
SABER_FULL(8, 48)
SABER_FULL(16, 32)
SABER_FULL(16, 40)
SABER_FULL(16, 48)
SABER_FULL(32, 32)
SABER_FULL(32, 40)
SABER_FULL(32, 48)
SABER_FULL(64, 32)
SABER_FULL(64, 40)
SABER_FULL(64, 48)
};

const FullVariant* full_variants_part_e(int* count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace saber
