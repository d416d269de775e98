// Ungapped X-drop extension of seed hits (xdrop_cat, src/search/engine/scan.rs:859-904;
// identical results to ungapped_xdrop_prof, src/search/engine.rs:2425-2455).
//
// For a seed (qp, sp) the right extension scores the pairs (qp+t, sp+t),
// t = 0..nr-1 (the seed pair included), the left extension (qp-t, sp-t),
// t = 1..nl. Each keeps a running sum, its best and the step of the first
// strict improvement, and stops after the step that falls more than x below
// the best. The result is (best_r + best_l, qp - st_t, qp + end_t).
//
// Work items are (hit, direction). The kernel is persistent: each lane runs
// one item at a time; whenever lanes finish, the wave claims that many new
// items with one atomic add and hands them out in lane order (mbcnt), so a
// wave never waits for its longest extension while other lanes idle. Scores
// come from the 32 x 32 int16 table in LDS (same as the band kernel) plus the
// query position's bias.
#include "common.h"

namespace saber {


__global__ __launch_bounds__(256) void xdrop_kernel(XdropArgs a) {
    __shared__ short T[kTableDim * kTableDim];
    for (int t = threadIdx.x; t < kTableDim * kTableDim; t += blockDim.x) T[t] = a.table[t];
    __syncthreads();
    const uint64_t nitems = 2ull * a.nhits;
    const uint32_t lane = threadIdx.x & 63;

    bool busy = false;
    uint32_t item = 0;
    const uint8_t* q = nullptr;
    const int8_t* b = nullptr;
    const uint8_t* s = nullptr;
    int64_t qpos = 0, spos = 0;
    int step = 1;               // +1 right, -1 left
    uint32_t n = 0, t = 0;      // steps available / taken
    int32_t cur = 0, best = 0;
    uint32_t bt = 0;
    bool out_of_work = false;

    while (true) {
        // Refill idle lanes from the global counter (one atomic per wave).
        const uint64_t idle = __ballot(!busy);
        if (idle && !out_of_work) {
            const uint32_t want = __popcll(idle);
            uint32_t base = 0;
            if (lane == (uint32_t)__ffsll((long long)idle) - 1) base = atomicAdd(a.next_item, want);
            base = __shfl(base, __ffsll((long long)idle) - 1);
            if (base >= nitems) out_of_work = true;
            if (!busy) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const uint64_t it = (uint64_t)base + rank;
                if (it < nitems) {
                    item = (uint32_t)it;
                    const uint32_t h = item >> 1;
                    const uint32_t qi = a.hits[4 * h + 0], si = a.hits[4 * h + 1];
                    const uint32_t qp = a.hits[4 * h + 2], sp = a.hits[4 * h + 3];
                    const uint64_t q0 = a.qoff[qi], s0 = a.soff[si];
                    const uint32_t m = (uint32_t)(a.qoff[qi + 1] - q0);
                    const uint32_t sl = (uint32_t)(a.soff[si + 1] - s0);
                    q = a.qres + q0;
                    b = a.qbias ? a.qbias + q0 : nullptr;
                    s = a.sres + s0;
                    if ((item & 1) == 0) {
                        step = 1;
                        qpos = qp;
                        spos = sp;
                        n = min(m - qp, sl - sp);
                    } else {
                        step = -1;
                        qpos = (int64_t)qp - 1;
                        spos = (int64_t)sp - 1;
                        n = min(qp, sp);
                    }
                    t = 0;
                    cur = 0;
                    best = 0;
                    bt = 0;
                    busy = n > 0;
                    if (!busy) {
                        a.part_best[item] = 0;
                        a.part_t[item] = 0;
                        a.part_steps[item] = 0;
                    }
                }
            }
        }
        if (!__ballot(busy)) {
            if (out_of_work) break;
            continue;
        }
        if (busy) {
            // Up to 8 steps per visit; the loop is the lane's own.
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t qc = q[qpos];
                const uint32_t sc = s[spos];
                const uint32_t qa = qc < 24 ? qc : (uint32_t)kCodeX;
                const uint32_t sa = sc < 24 ? sc : (uint32_t)kCodeX;
                int32_t v = T[qa * kTableDim + sa];
                if (b) v += b[qpos];
                cur += v;
                ++t;
                if (cur > best) {
                    best = cur;
                    bt = t;
                }
                const bool stop = (cur < best - a.x) || (t >= n);
                if (stop) {
                    a.part_best[item] = best;
                    a.part_t[item] = bt;
                    a.part_steps[item] = t;
                    busy = false;
                    break;
                }
                qpos += step;
                spos += step;
            }
        }
    }
}

__global__ void xdrop_combine_kernel(const uint32_t* hits, uint32_t nhits, const int32_t* pb,
                                     const uint32_t* pt, const uint32_t* ps, int32_t* score,
                                     uint32_t* q_lo, uint32_t* q_hi, uint32_t* steps) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhits) return;
    const uint32_t qp = hits[4 * h + 2];
    score[h] = pb[2 * h] + pb[2 * h + 1];
    q_lo[h] = qp - pt[2 * h + 1];
    q_hi[h] = qp + pt[2 * h];
    steps[h] = ps[2 * h] + ps[2 * h + 1];
}

hipError_t launch_xdrop(const XdropArgs& a, int32_t* score, uint32_t* q_lo, uint32_t* q_hi,
                        uint32_t* steps, int grid, hipStream_t s) {
    if (a.nhits == 0) return hipSuccess;
    hipLaunchKernelGGL(xdrop_kernel, dim3(grid), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t block = 256;
    hipLaunchKernelGGL(xdrop_combine_kernel, dim3((a.nhits + block - 1) / block), dim3(block), 0, s,
                       a.hits, a.nhits, a.part_best, a.part_t, a.part_steps, score, q_lo, q_hi, steps);
    return hipGetLastError();
}

}  // namespace saber
