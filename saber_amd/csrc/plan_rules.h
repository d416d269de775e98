// The window and band rules of plan_windows (Engine::window and
// plan_windows, src/search/engine.rs:1223-1324; Band, src/algorithm/banded.rs:30-54)
// for one region, shared by the host planner (host_pipeline.cpp, built by the
// host compiler) and the search's device pipeline (search_score.hip), so both
// apply the same rule.
#pragma once

#include <stdint.h>

#include "../../include/saber_hip.h"

#if defined(__HIPCC__)
#define SABER_HD __host__ __device__
#else
#define SABER_HD
#endif

namespace saber_plan {

constexpr uint64_t kBandScoreMinCells = 1ull << 20;  // engine.rs:401
constexpr uint64_t kBandCostRatio = 12;              // engine.rs:402
constexpr int64_t kBandBridge = 1024;                // engine.rs:404
constexpr int64_t kSlackUnbounded = INT64_MAX / 8;   // window_slack at or above: no band

SABER_HD inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
SABER_HD inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }

struct Band {
    int64_t lo, hi;
    SABER_HD int64_t width() const { return imax(hi - lo + 1, 0); }
    SABER_HD Band clip(int64_t m, int64_t n) const { return Band{imax(lo, -(m - 1)), imin(hi, n - 1)}; }
    SABER_HD bool covers(int64_t m, int64_t n) const { return lo <= -(m - 1) && hi >= n - 1; }
    SABER_HD Band widen(int64_t by) const { return Band{lo - by, hi + by}; }
};

// Engine::window (engine.rs:1223-1234): the subject window of diagonals
// [d_min, d_max] with the slack, or the whole subject when that covers at
// least three quarters of it.
SABER_HD inline void window(int64_t d_min, int64_t d_max, int64_t qlen, int64_t sl, int64_t slack, uint64_t* olo,
                            uint64_t* ohi) {
    const int64_t lo = imin(imax(d_min - slack, 0), sl);
    const int64_t hi = imax(imin(qlen + d_max + slack, sl), lo);
    if ((hi - lo) * 4 >= sl * 3) {
        *olo = 0;
        *ohi = (uint64_t)sl;
    } else {
        *olo = (uint64_t)lo;
        *ohi = (uint64_t)hi;
    }
}

// Regions of one subject at or above this best are strong and chain into a
// common band (plan_windows' strong set).
SABER_HD inline int32_t strong_cut(int32_t hsp_threshold, int32_t top) {
    const int32_t a = 2 * hsp_threshold, b = top / 4;
    return a > b ? a : b;
}

// The plan of one region of a query of length q against a subject of length
// sl, given the diagonal span [span_lo, span_hi] of its chain of strong
// regions (its own diagonals when it is not strong).
SABER_HD inline saber_hip_plan plan_region(int64_t d_min, int64_t d_max, int32_t best, int64_t span_lo,
                                           int64_t span_hi, int64_t q, int64_t sl, const saber_hip_plan_params& p) {
    saber_hip_plan o{};
    window(d_min, d_max, q, sl, p.window_slack, &o.lo, &o.hi);
    o.banded = 0;
    if (p.exhaustive || best == 0 || p.window_slack >= kSlackUnbounded) return o;
    const int64_t dmin = span_lo, dmax = span_hi;
    if (p.score_band > 0) {
        const int64_t w = p.score_band;
        const int64_t lo = imin(imax(dmin - w, 0), sl);
        const int64_t hi = imax(imin(q + dmax + w, sl), lo);
        const int64_t nn = hi - lo;
        const Band band = Band{dmin - w - lo, dmax + w - lo}.clip(q, nn);
        o.lo = (uint64_t)lo;
        o.hi = (uint64_t)hi;
        o.banded = (nn == 0 || band.covers(q, nn)) ? 0 : 1;
        o.band_lo = o.banded ? band.lo : 0;
        o.band_hi = o.banded ? band.hi : 0;
        return o;
    }
    if ((uint64_t)q * (o.hi - o.lo) < kBandScoreMinCells) return o;
    uint64_t ulo = 0, uhi = 0;
    window(dmin, dmax, q, sl, p.window_slack, &ulo, &uhi);
    const int64_t nn = (int64_t)(uhi - ulo);
    const Band band = Band{dmin - (int64_t)ulo, dmax - (int64_t)ulo}.widen(p.window_slack).clip(q, nn);
    if (band.covers(q, nn)) return o;
    const uint64_t rows = (uint64_t)imin(q, nn + band.width());
    const uint64_t rect = (uint64_t)q * (uint64_t)nn;
    if ((uint64_t)band.width() * rows * kBandCostRatio < rect) {
        o.lo = ulo;
        o.hi = uhi;
        o.banded = 1;
        o.band_lo = band.lo;
        o.band_hi = band.hi;
    }
    return o;
}

}  // namespace saber_plan
