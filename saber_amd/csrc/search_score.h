// The device-resident candidate-scoring stage of the search (search_score.hip):
// from the selected regions of region_kernel to each region's exact band
// score, without the regions leaving the device. Per query the cap under
// region_order (scan.rs:1453-1466, engine.rs:493-495), plan_windows
// (engine.rs:1247-1324), the two-stage cut (score_windows, engine.rs:1021-1099)
// and the optional learned prefilter; then only the regions that can pass the
// reporting cut are copied back for the host's finish (engine.rs:1127-1221).
#pragma once

#include <vector>

#include "ctx.h"

namespace saber {

// A region that may be reported: its index in the capped list, the region,
// plan, exact score and subject length.
struct SearchCand {
    uint64_t ri;
    saber_hip_region r;
    saber_hip_plan p;
    int32_t score;
    uint32_t q;
    uint64_t slen;
};

struct SearchScoreIn {
    hipStream_t st;
    const saber_hip_region* regions;   // device: region_kernel's selected regions, key order; reserved = query
    uint64_t nregions;
    uint32_t nq;
    const uint64_t* d_qoff;            // device: query offsets (nq + 1)
    const uint64_t* d_qlen;            // device: query lengths
    const uint8_t* d_qres;             // device: query codes (data start, guarded)
    uint64_t qtotal;
    const int8_t* d_qbias;             // device: composition biases (data start, guarded) or null
    const saber_hip_query_stats* qstats;  // host, per query
    const int64_t* floor;              // host, per query: min_score at the E-value cut
    const int64_t* region_floor;       // host, per query: min_score at the region E-value
    const int32_t* region_thr;         // host, per query: the region threshold
    double region_e, ecut;
    const saber_hip_search_params* P;
    const saber_hip_scheme* scoring;
    const saber_hip_prefilter* prefilter;  // may be null
    bool keep_regions;
};

struct SearchScoreOut {
    std::vector<uint64_t> roff;        // nq + 1: capped regions per query
    std::vector<SearchCand> cands;     // regions that can pass the reporting cut, in capped order
    uint64_t regions = 0, dropped_by_cap = 0, band_scored = 0, rescored = 0, banded_cells = 0, pruned = 0;
    double ms_regions = 0, ms_plan = 0, ms_prefilter = 0, ms_score = 0;
    // keep_regions: every capped region, in capped order
    std::vector<saber_hip_region> flat;
    std::vector<saber_hip_plan> plans;
    std::vector<int32_t> scores;
    std::vector<uint8_t> rescored_flag, pruned_flag;
    std::vector<uint32_t> qidx;
    std::vector<uint64_t> slen;
};

// Runs the stage on the context stream; the caller holds c->mu.
int search_score_device(saber_hip_ctx* c, const SearchScoreIn& in, SearchScoreOut* out);

}  // namespace saber
