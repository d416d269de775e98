// Exact int32 banded Smith-Waterman kernel: one thread per job.
//
// The rescue path of the packed kernel (jobs whose int16 best reached 32767,
// bands wider than the packed variants, unusual scoring parameters) and the
// reference for it in tests. It follows the scalar recurrence of
// banded_score_scalar (src/algorithm/banded.rs:134-224) literally, with the
// two band rows in a per-job slice of a device work buffer.
#include "common.h"

namespace saber {

constexpr int32_t kNeg32 = INT32_MIN / 4;

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

struct JobResult {
    int32_t score;
    uint32_t q_end, s_end;
    uint64_t cells;
};

// One job, its two band rows in work (4 * (w + 1) int32).
__device__ JobResult generic_run(const GenericArgs& a, const BandJobDev& J, int32_t* work) {
    const int64_t w = J.w;
    int32_t* hp = work;
    int32_t* fp = hp + (w + 1);
    int32_t* hc = fp + (w + 1);
    int32_t* fc = hc + (w + 1);
    for (int64_t k = 0; k <= w; ++k) {
        hp[k] = kNeg32;
        fp[k] = kNeg32;
        hc[k] = kNeg32;
        fc[k] = kNeg32;
    }
    const uint8_t* q = (J.swapped ? a.sres : a.qres) + J.q_off;
    const uint8_t* s = (J.swapped ? a.qres : a.sres) + J.s_off;
    const int8_t* bias = a.qbias && !J.swapped ? a.qbias + J.q_off : nullptr;
    const int64_t m = J.m, n = J.n, lo = J.lo;
    const int32_t go = a.go, ge = a.ge;
    int32_t best = 0;
    int64_t bq = 0, bs = 0;
    uint64_t cells = 0;
    // Rows before i_first hold no band cell: all of their entries are -inf in
    // the scalar kernel, which is the initial state here.
    const int64_t i_end = (int64_t)J.i_first + J.rows;
    for (int64_t i = J.i_first; i < i_end; ++i) {
        int64_t kmin = -(i + lo);
        if (kmin < 0) kmin = 0;
        int64_t kmax = n - 1 - i - lo;
        if (kmax > w - 1) kmax = w - 1;
        const uint32_t qc = q[i] < 24 ? q[i] : (uint32_t)kCodeX;
        const int32_t* row = a.mat + qc * 24;
        const int32_t b = bias ? (int32_t)bias[i] : 0;
        int32_t e = kNeg32, hl = kNeg32;
        for (int64_t k = 0; k < kmin; ++k) {
            hc[k] = kNeg32;
            fc[k] = kNeg32;
        }
        for (int64_t k = kmin; k <= kmax; ++k) {
            const int64_t j = i + lo + k;
            const int32_t ecand = imax(hl - go, e - ge);
            e = k > kmin ? ecand : kNeg32;
            const int32_t f = i >= 1 ? imax(hp[k + 1] - go, fp[k + 1] - ge) : kNeg32;
            const int32_t hd = (i >= 1 && j >= 1) ? hp[k] : 0;
            const uint32_t sc = s[j] < 24 ? s[j] : (uint32_t)kCodeX;
            const int32_t diag = hd + row[sc] + b;
            const int32_t h = imax(imax(diag, e), imax(f, 0));
            hc[k] = h;
            fc[k] = f;
            hl = h;
            if (h > best) {
                best = h;
                bq = i + 1;
                bs = j + 1;
            }
        }
        for (int64_t k = kmax + 1; k < w; ++k) {
            hc[k] = kNeg32;
            fc[k] = kNeg32;
        }
        cells += (uint64_t)(kmax - kmin + 1);
        int32_t* x = hp;
        hp = hc;
        hc = x;
        x = fp;
        fp = fc;
        fc = x;
    }
    return JobResult{best, best > 0 ? (uint32_t)bq : 0u, best > 0 ? (uint32_t)bs : 0u, cells};
}

__device__ void generic_job(const GenericArgs& a, uint32_t id, int32_t* work) {
    const JobResult r = generic_run(a, a.jobs[id], work);
    a.out.score[id] = r.score;
    a.out.q_end[id] = r.q_end;
    a.out.s_end[id] = r.s_end;
    a.out.cells[id] = r.cells;
}

// One thread per listed job, work at work_off[thread].
__global__ __launch_bounds__(128) void generic_band_kernel(GenericArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nids) return;
    generic_job(a, a.ids[t], a.work + a.work_off[t]);
}

// The jobs a.ids[0 .. *count) for a count known only on the device (the
// packed kernel's saturated jobs): a fixed grid walks the list, each thread
// with its own work slot of slot_ints int32.
__global__ __launch_bounds__(128) void generic_slots_kernel(GenericArgs a, const uint32_t* count, uint32_t slot_ints) {
    const uint32_t n = *count;
    const uint32_t nt = gridDim.x * blockDim.x;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t* work = a.work + (uint64_t)t * slot_ints;
    for (uint32_t i = t; i < n; i += nt) generic_job(a, a.ids[i], work);
}

hipError_t launch_generic_slots(const GenericArgs& a, const uint32_t* count, uint32_t nthreads, uint32_t slot_ints,
                                hipStream_t s) {
    if (nthreads == 0) return hipSuccess;
    const uint32_t block = 128;
    SABER_LAUNCH(generic_slots_kernel, dim3((nthreads + block - 1) / block), dim3(block), 0, s, a, count,
                       slot_ints);
    return hipGetLastError();
}

// Full-matrix pairs (query, subject) = (pairs[2i], pairs[2i+1]) for
// i < *count, scored exactly: the whole matrix as one band (lo = -(m-1),
// width m + n - 1), results written to row query, column subject of the
// [query][subject] outputs (stride nsub). A fixed grid walks the list, each
// thread with its own work slot of slot_ints int32.
__global__ __launch_bounds__(128) void full_int32_kernel(GenericArgs a, const uint32_t* pairs, const uint32_t* count,
                                                         const uint64_t* qoff, const uint64_t* soff, uint32_t nsub,
                                                         int32_t* score, uint32_t* q_end, uint32_t* s_end,
                                                         uint32_t slot_ints) {
    const uint32_t n = *count;
    const uint32_t nt = gridDim.x * blockDim.x;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t* work = a.work + (uint64_t)t * slot_ints;
    for (uint32_t i = t; i < n; i += nt) {
        const uint32_t qi = pairs[2 * i], si = pairs[2 * i + 1];
        const int64_t m = (int64_t)(qoff[qi + 1] - qoff[qi]), len = (int64_t)(soff[si + 1] - soff[si]);
        BandJobDev J{};
        J.q_off = qoff[qi];
        J.s_off = soff[si];
        J.m = (uint32_t)m;
        J.n = (uint32_t)len;
        J.lo = (int32_t)(m > 0 ? -(m - 1) : 0);
        J.w = (int32_t)(m > 0 && len > 0 ? m + len - 1 : 0);
        J.i_first = 0;
        J.rows = J.w > 0 ? (int32_t)m : 0;
        const JobResult r = generic_run(a, J, work);
        const uint64_t o = (uint64_t)qi * nsub + si;
        score[o] = r.score;
        q_end[o] = r.q_end;
        s_end[o] = r.s_end;
    }
}

hipError_t launch_full_int32(const GenericArgs& a, const uint32_t* pairs, const uint32_t* count, const uint64_t* qoff,
                             const uint64_t* soff, uint32_t nsub, int32_t* score, uint32_t* q_end, uint32_t* s_end,
                             uint32_t nthreads, uint32_t slot_ints, hipStream_t s) {
    if (nthreads == 0) return hipSuccess;
    const uint32_t block = 128;
    SABER_LAUNCH(full_int32_kernel, dim3((nthreads + block - 1) / block), dim3(block), 0, s, a, pairs, count,
                       qoff, soff, nsub, score, q_end, s_end, slot_ints);
    return hipGetLastError();
}

hipError_t launch_generic(const GenericArgs& a, hipStream_t s) {
    if (a.nids == 0) return hipSuccess;
    const uint32_t block = 128;
    const uint32_t grid = (a.nids + block - 1) / block;
    SABER_LAUNCH(generic_band_kernel, dim3(grid), dim3(block), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
