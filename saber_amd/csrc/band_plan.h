// Band batch planning shared by the host planner (api.hip) and the device
// planner (band_plan.hip): the planned geometry of a job, the cost model of
// the latency rule and the device planner's interface.
#pragma once

#include "ctx.h"

namespace saber {

// One lane steps through every row of its job, so a job's time is at least
// rows x the lane's per-row time. Jobs whose time at their throughput shape
// would exceed this fraction of the batch's estimated time are planned for
// low per-row latency instead (short lanes, launched first). A context can
// replace the rule by a fixed row count (saber_hip_band_tune).
constexpr double kLatencyShare = 0.6;

// Rates for the band planner's time estimates (lane VALU instructions per
// ms): the whole GPU at about half its issue rate, and one lane of a wave
// that shares its SIMD with another.
constexpr double kGpuLaneInstrPerMs = 3.4e10;
constexpr double kLaneInstrPerMs = 3.0e5;

// Query / window lengths the packed band kernel accepts (window columns are
// handled as 8 * j in 32 bits).
constexpr int64_t kMaxPackedLen = int64_t(1) << 27;

// Batches with at least this many jobs are planned on the device unless the
// context asks for one planner (SABER_HIP_HOST_PLAN / SABER_HIP_DEVICE_PLAN).
constexpr uint64_t kDevicePlanMinJobs = 1u << 16;

// sum of x for x in [a, b] (0 when b < a)
__host__ __device__ inline int64_t range_sum(int64_t a, int64_t b) { return b < a ? 0 : (a + b) * (b - a + 1) / 2; }

// Band cells of a planned job: every row i in [i_first, i_first + rows) holds
// the columns max(0, i + lo) .. min(n - 1, i + lo + w - 1), at least one
// (the planner keeps exactly those rows), so the count is a sum of clamped
// linear terms.
__host__ __device__ inline uint64_t band_cells(const BandJobDev& d) {
    if (d.rows <= 0) return 0;
    const int64_t i0 = d.i_first, i1 = (int64_t)d.i_first + d.rows - 1;
    const int64_t lo = d.lo, hi = (int64_t)d.lo + d.w - 1, n = d.n;
    // sum of min(n - 1, i + hi): i + hi up to i = n - 1 - hi, then n - 1
    const int64_t ic = n - 1 - hi;
    const int64_t i1c = i1 < ic ? i1 : ic;
    int64_t right = range_sum(i0 + hi, i1c + hi);
    const int64_t from = i0 > ic + 1 ? i0 : ic + 1;
    right += (i1 - from + 1 > 0 ? i1 - from + 1 : 0) * (n - 1);
    // sum of max(0, i + lo): 0 below i = -lo
    const int64_t left = range_sum((i0 > -lo ? i0 : -lo) + lo, i1 + lo);
    return (uint64_t)(right - left + (i1 - i0 + 1));
}

// The planner's cost model of one shape: lanes per job and the VALU
// instructions of one lane per row (Shape::row_cost, an integer).
struct PlanShape {
    int32_t L;
    uint32_t row_cost;
};

// The jobs of a batch as the planners take them. The offsets are host arrays
// (count + 1 entries); soffs_dev, when set, is the same subject offsets
// already on the device (the context database). jobs_dev / qoffs_dev, when
// set, are the jobs and query offsets already on the device (then only the
// device planner can take the batch, and jobs / qoffs may be null).
struct BandPlanInput {
    const saber_hip_band_job* jobs;
    uint64_t njobs;
    const uint64_t* qoffs;
    uint32_t qcount;
    const uint64_t* soffs;
    const uint64_t* soffs_dev;
    uint32_t nsub;
    bool packed_ok, swap_ok;
    const saber_hip_band_job* jobs_dev = nullptr;
    const uint64_t* qoffs_dev = nullptr;
    bool ends = false;   // end cells: the shape choice of end-cell batches
};

// A plan as the band batch runs it. Device buffers: jobs and cells (one per
// job), pairs (2 job ids per pair, kNone = none), blocks (chunks in dispatch
// order), gen_ids / gen_off (the jobs for the int32 kernel, in job order,
// and their work offsets); counts on the host.
struct BandPlan {
    uint64_t npairs = 0, nchunks = 0;
    uint32_t ngeneric = 0;
    uint64_t npacked = 0;   // jobs of the packed kernel (the rest of the batch
                            // beyond npacked + ngeneric has no band cells)
    uint64_t gen_work = 0;  // int32 work elements of the generic jobs
    int32_t wmax = 0;       // widest packed job (0: none)
};

// Plans the batch on the device (band_plan.hip), writing the buffers above.
// Fails with SABER_HIP_EINVAL and the host planner's messages for bad jobs.
int band_plan_device(saber_hip_ctx* c, hipStream_t st, const BandPlanInput& in, DevBuf<BandJobDev>& jobs,
                     DevBuf<uint64_t>& cells, DevBuf<uint32_t>& pairs, DevBuf<BandBlock>& blocks,
                     DevBuf<uint32_t>& gen_ids, DevBuf<uint64_t>& gen_off, BandPlan* out);

}  // namespace saber
