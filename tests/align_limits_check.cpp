// Host-only check of the traceback lane kernel's end-cell key limit
// (saber_amd/csrc/align_limits.h): every key of a job routed to a lane class
// is below the none value 0xFFFFFFFF, and the first row count that would
// reach it is refused. Prints "ok" or the failing case.
#include <cstdio>

#include "../saber_amd/csrc/align_limits.h"

using namespace saber;

int main() {
    for (int x = 0; x < kAlignLaneClasses; ++x) {
        const uint64_t W = (uint64_t)align_class_lanes(x) * align_class_cells(x);
        // largest row count whose largest key (rows - 1) * W + W - 1 stays below 0xFFFFFFFF
        const uint64_t max_rows = (0xFFFFFFFFull - 1) / W;
        if (!align_lane_key_fits(max_rows, x) || align_lane_key_fits(max_rows + 1, x)) {
            std::printf("class %d (W %llu): limit not at %llu rows\n", x, (unsigned long long)W,
                        (unsigned long long)max_rows);
            return 1;
        }
        if ((max_rows - 1) * W + (W - 1) >= 0xFFFFFFFFull) {
            std::printf("class %d: largest key reaches the none value\n", x);
            return 1;
        }
        if (!align_lane_key_fits(1, x) || !align_lane_key_fits(32000, x)) {
            std::printf("class %d: ordinary job refused\n", x);
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
