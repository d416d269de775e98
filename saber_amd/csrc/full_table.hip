// Concatenates the compiled full-matrix kernel variants into one table.
#include <vector>

#include "common.h"

namespace saber {

const FullVariant* full_variants_part_e(int* count);
const FullVariant* full_variants_part_f(int* count);
const FullVariant* full_variants_part_g(int* count);
const FullVariant* full_variants_part_h(int* count);

const FullVariant* full_variants(int* count) {
    static const std::vector<FullVariant> table = [] {
        std::vector<FullVariant> t;
        for (auto f : {full_variants_part_e, full_variants_part_f, full_variants_part_g, full_variants_part_h}) {
            int n = 0;
            const FullVariant* p = f(&n);
            t.insert(t.end(), p, p + n);
        }
        return t;
    }();
    *count = (int)table.size();
    return table.data();
}

}  // namespace saber
