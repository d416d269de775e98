// Concatenates the compiled band kernel variants into one table.
#include <vector>

#include "common.h"

namespace saber {

const BandVariant* band_variants_part_a(int* count);
const BandVariant* band_variants_part_b(int* count);
const BandVariant* band_variants_part_c(int* count);
const BandVariant* band_variants_part_d(int* count);

const BandVariant* band_variants(int* count) {
    static const std::vector<BandVariant> table = [] {
        std::vector<BandVariant> t;
        for (auto f : {band_variants_part_a, band_variants_part_b, band_variants_part_c,
                       band_variants_part_d}) {
            int n = 0;
            const BandVariant* p = f(&n);
            t.insert(t.end(), p, p + n);
        }
        return t;
    }();
    *count = (int)table.size();
    return table.data();
}

}  // namespace saber
