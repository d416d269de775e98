// Lane classes of the banded traceback kernel and the limit of its packed
// end-cell key. Plain constexpr functions (host and device alike), with no
// dependence on the HIP headers, so host-only code and tests can use them.
#pragma once
#include <stdint.h>

namespace saber {

// Lane classes: x < 7: 1 << x lanes of 17 cells (band widths up to
// 17 << x: the widths 2^k + 1 of half-width bands fit); x >= 7:
// 4 << (x - 7) lanes of 5 cells (widths up to 20 << (x - 7)), for the jobs
// that would otherwise run alone at the end of a launch: a quarter of the
// cells per lane makes each row step about a quarter as long, at a somewhat
// higher cost per cell.
constexpr int kAlignLaneClasses = 12, kAlignWideClasses = 7;
constexpr uint64_t kAlignMinLongSteps = 64;   // jobs of fewer row steps keep 17 cells a lane
constexpr int align_class_lanes(int x) { return x < kAlignWideClasses ? 1 << x : 4 << (x - kAlignWideClasses); }
constexpr int align_class_cells(int x) { return x < kAlignWideClasses ? 17 : 5; }

// The lane kernel packs a job's end cell as (row - i_first) * W + k in 32
// bits, W = lanes x cells of its class, with 0xFFFFFFFF meaning none: a job
// of `rows` rows fits a class when rows * W < 0xFFFFFFFF (every key of its
// rows is below the none value). Jobs that fit no class run on one thread.
constexpr bool align_lane_key_fits(uint64_t rows, int x) {
    return rows * (uint64_t)(align_class_lanes(x) * align_class_cells(x)) < 0xFFFFFFFFull;
}

}  // namespace saber
