// Host-side internals shared by the library's translation units: the
// context, device buffers, error reporting and the upload helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/saber_hip.h"
#include "common.h"

namespace saber {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string& msg);
const char* g_err_cstr();

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::saber::fail(SABER_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Device memory of the library's buffers (dev_cache.hip): blocks released
// with dev_free are cached per device after the device has finished its
// work, and handed out again by dev_alloc.
hipError_t dev_alloc(void** p, uint64_t bytes);
void dev_free(void* p);

template <class T>
struct DevBuf {
    T* p = nullptr;
    uint64_t n = 0;
    hipError_t alloc(uint64_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return dev_alloc(reinterpret_cast<void**>(&p), sizeof(T) * count);
    }
    // Keeps the allocation when it already holds count elements.
    hipError_t reserve(uint64_t count) {
        if (p && n >= count) return hipSuccess;
        return alloc(count < 1 ? 1 : count);
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        n = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
};

struct Shape {
    int L, C, K;
    int W() const { return L * C; }
    // Every lane caps its last K cells, so a shape serves widths W-K+1 .. W,
    // and every width up to W when K = C.
    bool covers(int64_t w) const { return w >= 1 && w <= W() && (K == C || w > W() - K); }
    // The end-cell build keeps a register copy of a lane's cells of the row
    // where its best last rose; lanes of more than 21 cells would spill it.
    bool ends_ok() const { return C <= 21; }
    // VALU instructions of one lane per row: about 13 per packed cell, one per
    // capped cell and a fixed part for the row and stream bookkeeping.
    double row_cost() const { return 13.0 * C + K + 45.0; }
    // Work of one job pair per row.
    double pair_cost() const { return L * row_cost(); }
};

// Residue and bias buffers carry guard bytes so kernels may read a few bytes
// before and after any sequence (clamped into [-16, len + 64)).
constexpr uint64_t kGuard = 16, kGuardTail = 64;

hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
// SABER_HIP_EDEVICE when the context injects faults (SABER_HIP_INJECT_FAULTS).
int injected(const saber_hip_ctx* c);
int check_seqset(const saber_hip_seqset* s, const char* what);

template <class T>
int upload(DevBuf<T>& d, const T* h, uint64_t n, hipStream_t s) {
    HIP_TRY(d.alloc(n == 0 ? 1 : n));
    if (n) HIP_TRY(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return SABER_HIP_OK;
}

// upload() into an allocation kept when it is large enough (buffers that a
// caller refills call after call).
template <class T>
int upload_keep(DevBuf<T>& d, const T* h, uint64_t n, hipStream_t s) {
    HIP_TRY(d.reserve(n == 0 ? 1 : n));
    if (n) HIP_TRY(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return SABER_HIP_OK;
}

template <class T>
int upload_guarded(DevBuf<T>& d, const T* h, uint64_t n, hipStream_t s) {
    HIP_TRY(d.alloc(n + kGuard + kGuardTail));
    HIP_TRY(hipMemsetAsync(d.p, 0, sizeof(T) * (n + kGuard + kGuardTail), s));
    if (n) HIP_TRY(hipMemcpyAsync(d.p + kGuard, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return SABER_HIP_OK;
}

// A residue buffer: guarded upload, then every code >= 24 stored as X.
int upload_codes(DevBuf<uint8_t>& d, const uint8_t* h, uint64_t n, hipStream_t s);

// Seed-scan data of the uploaded database (search.hip), built on first use
// for a seed scheme and kept until the database changes.
struct ScanDbCache;
void scan_cache_free(ScanDbCache* p);
struct ScanCacheDeleter {
    void operator()(ScanDbCache* p) const { scan_cache_free(p); }
};

// Device scratch of the band planner (band_plan.hip), kept across batches
// and grown to the largest.
struct BandPlanScratch;
void plan_scratch_free(BandPlanScratch* p);
struct PlanScratchDeleter {
    void operator()(BandPlanScratch* p) const { plan_scratch_free(p); }
};

// The banded part of score_windows (saber_hip_score_regions) with the band
// cells of the banded plans' passes summed into *banded_cells (may be null)
// and regions with skip[i] != 0 (may be null) left unscored (score 0).
int score_regions_impl(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                       const uint32_t* qidx, const saber_hip_region* regs, const saber_hip_plan* plans, uint64_t n,
                       const saber_hip_scheme* scheme, int64_t narrow_band, int64_t full_band, const double* thr,
                       int32_t* out, uint8_t* rescored, uint64_t* banded_cells, const uint8_t* skip);

// Band jobs, query offsets (count + 1), codes and biases (data starts of
// guarded buffers; qbias may be null) already in device memory; subjects are
// the context database.
struct DevBandInput {
    const saber_hip_band_job* jobs;
    const uint64_t* qoffs;
    uint32_t qcount;
    const uint8_t* qres;
    uint64_t qlen_total;
    const int8_t* qbias;
};
// saber_hip_band_batch_create over host inputs (dev == null) or device inputs
// (queries == null, subjects == null). The caller holds c->mu.
int band_batch_create_impl(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                           const saber_hip_band_job* jobs, uint64_t njobs, const DevBandInput* dev,
                           const saber_hip_scheme* scheme, int ends, saber_hip_band_batch** out);
// One scoring pass of a batch enqueued on the context stream (the caller
// holds c->mu); its per-job scores and band cells stay in device memory.
int band_batch_enqueue(saber_hip_band_batch* b);
const int32_t* band_batch_scores(const saber_hip_band_batch* b);
const uint64_t* band_batch_cells(const saber_hip_band_batch* b);

// Takes the address of saber_hip_xdrop_noalloc's scratch array on the
// current device (xdrop_noalloc.hip); called when a context is created.
int xdrop_noalloc_setup(int device);

// Prefilter MLP forward over device features (bf16 [n][64]) into device
// logits, enqueued on st; the model's context (its device).
int mlp_forward_device(saber_hip_mlp* m, const __bf16* X, uint64_t n, float* logits, hipStream_t st);
saber_hip_ctx* mlp_context(const saber_hip_mlp* m);

// Full-SW launches of one run are spread over this many streams (all at the
// default priority) so small subject buckets and the tails of large ones
// overlap.
constexpr int kStreams = 4;

}  // namespace saber

struct saber_hip_ctx {
    int device = 0;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;
    hipStream_t work[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex mu;
    saber::DevBuf<uint8_t> db;
    std::vector<uint64_t> db_off;
    saber::DevBuf<uint64_t> db_off_dev;   // the same offsets on the device
    std::vector<uint8_t> db_host;         // host copy of the residues (canonical codes)
    uint64_t db_gen = 0;                  // incremented by every upload_db
    int cus = 0;                  // compute units of the device
    // tuned full-matrix shapes: default shape id -> chosen shape id, by
    // flavour (0 scores only, 1 end cells)
    std::map<int, int> full_tuned[2];
    std::vector<int> choice;      // width -> shape index (-1: exact int32 kernel)
    std::vector<int> choice_lat;  // the same for jobs with many rows
    std::vector<int> choice_ends, choice_lat_ends;  // the same for end-cell batches
    std::vector<saber::Shape> shapes;
    int latency_rows = -1;  // -1: kLatencyShare rule; 0: off; else a row count
    std::unique_ptr<saber::ScanDbCache, saber::ScanCacheDeleter> scan;
    std::unique_ptr<saber::BandPlanScratch, saber::PlanScratchDeleter> plan;
};
