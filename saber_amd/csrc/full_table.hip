// Shape table and dispatch of the persistent full-matrix kernels.
#include <vector>

#include "full_kernel.h"

namespace saber {

hipError_t launch_full_multi_nobias_scores(int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_full_multi_bias_scores(int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_full_multi_nobias_ends(int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_full_multi_bias_ends(int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s);
int full_multi_blocks_nobias_scores(int nt);
int full_multi_blocks_bias_scores(int nt);
int full_multi_blocks_nobias_ends(int nt);
int full_multi_blocks_bias_ends(int nt);

const FullShape* full_shapes(bool ends, int* count) {
#define SABER_FULL_SCORE_ROW(L_, C_) FullShape{L_, C_, kFullThreads<C_, false>},
#define SABER_FULL_END_ROW(L_, C_) FullShape{L_, C_, kFullThreads<C_, true>},
    static const std::vector<FullShape> scores = {SABER_FULL_SCORE_SHAPES(SABER_FULL_SCORE_ROW)};
    static const std::vector<FullShape> endc = {SABER_FULL_END_SHAPES(SABER_FULL_END_ROW)};
#undef SABER_FULL_SCORE_ROW
#undef SABER_FULL_END_ROW
    const std::vector<FullShape>& t = ends ? endc : scores;
    *count = (int)t.size();
    return t.data();
}

hipError_t launch_full_multi(bool bias, bool ends, int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s) {
    if (ends) return bias ? launch_full_multi_bias_ends(nt, a, grid, s) : launch_full_multi_nobias_ends(nt, a, grid, s);
    return bias ? launch_full_multi_bias_scores(nt, a, grid, s) : launch_full_multi_nobias_scores(nt, a, grid, s);
}

int full_multi_blocks_per_cu(bool bias, bool ends, int nt) {
    if (ends) return bias ? full_multi_blocks_bias_ends(nt) : full_multi_blocks_nobias_ends(nt);
    return bias ? full_multi_blocks_bias_scores(nt) : full_multi_blocks_nobias_scores(nt);
}

}  // namespace saber
