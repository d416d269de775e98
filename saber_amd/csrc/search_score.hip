// The search's candidate-scoring stage on the device (see search_score.h):
// region_kernel's selected regions stay in device memory through the
// per-query cap, plan_windows, the two-stage thresholds, the optional learned
// prefilter and both band-scoring passes; the host receives only the regions
// that can pass the reporting cut.
//
// Each step restates the host rule it replaces:
//  * cap: a stable radix sort by (query, best descending) of the regions in
//    region_kernel's order (query, subject, best descending, d_min) gives
//    region_order (best desc, subject, d_min, d_max; engine.rs:493-495) within
//    each query; the first max_candidates of each query are kept
//    (scan.rs:1453-1466).
//  * plans: the capped regions grouped by (query, subject) in capped order;
//    one thread per group forms the chains of strong regions (plan_windows,
//    engine.rs:1247-1324: strong = best >= max(2 hsp_threshold, top / 4),
//    sorted by (d_min, d_max, position), single linkage within 1024
//    diagonals); each region's plan is plan_rules.h's rule, the host planner's.
//  * thresholds: score_prefilter_frac x floor_lower_bound (engine.rs:1057-1068,
//    stats.rs:700-716) in double precision; the device's log may differ from
//    the host's in the last bits, so a threshold (or the ceil inside the
//    search-space bound) within 1e-9 of an integer is recomputed by the host.
//  * scores: score_windows' narrow pass, cut and full pass (engine.rs:1021-1099)
//    as two device-planned band batches over device job lists.
//  * candidates: the host's reporting loop reads only regions whose score can
//    pass the E-value cut; the device keeps every region whose E-value lower
//    bound (stats.rs:725-733) is within 1e-6 of the cut or below (a superset),
//    and the host applies the exact test.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "plan_rules.h"
#include "search_score.h"

namespace saber {

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

constexpr uint32_t kT = 256;
constexpr uint32_t kMaxGroup = 8;  // regions of one (query, subject): at most max_regions_per_subject <= 8

// Every kernel here loops over its range with the grid's stride: 2048
// workgroups fill the device once and keep the per-wave counter updates of
// the counting kernels few (updates of one address serialise).
uint32_t grid_of(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kT - 1) / kT, 2048)); }
int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __host__ inline bool near_integer(double x) { return fabs(x - rint(x)) <= 1e-9 * fmax(1.0, fabs(x)); }

// space_lower_bound (stats.rs:725-733); *amb when its ceil is within 1e-9 of
// an integer.
__device__ inline double space_lb(const saber_hip_query_stats& q, uint64_t slen, bool* amb) {
    if (!q.per_subject) return q.search_space;
    const double m = (double)(q.query_len > 0 ? q.query_len : 1);
    const double n = (double)(slen ? slen : 1);
    const double a = fmax(q.alpha_d_lambda * (q.log_k + log(m * n)) + q.beta, 0.0);
    if (near_integer(a)) *amb = true;
    const double up = ceil(a);
    return q.num_seqs * fmax(m - up, 1.0) * fmax(n - up, 1.0);
}

// ---- cap ----

__global__ __launch_bounds__(kT) void cap_keys_kernel(const saber_hip_region* r, uint64_t n, uint64_t* keys,
                                                      uint32_t* idx, uint32_t* qcount) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        const uint32_t q = r[i].reserved;
        const int32_t b = r[i].best < 0 ? 0 : r[i].best;
        keys[i] = ((uint64_t)q << 31) | (uint64_t)(0x7FFFFFFFu - (uint32_t)b);
        idx[i] = (uint32_t)i;
        atomicAdd(&qcount[q], 1u);
    }
}

__global__ __launch_bounds__(kT) void cap_scatter_kernel(const uint64_t* skeys, const uint32_t* sidx, uint64_t n,
                                                         const saber_hip_region* r, const uint64_t* qstart,
                                                         const uint64_t* roff, uint64_t cap, saber_hip_region* flat,
                                                         uint32_t* qidx) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        const uint32_t q = (uint32_t)(skeys[i] >> 31);
        const uint64_t rank = i - qstart[q];
        if (rank < cap) {
            flat[roff[q] + rank] = r[sidx[i]];
            qidx[roff[q] + rank] = q;
        }
    }
}

// ---- plans ----

__global__ __launch_bounds__(kT) void group_keys_kernel(const saber_hip_region* flat, const uint32_t* qidx,
                                                        uint64_t n, uint64_t* keys, uint32_t* idx) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        keys[i] = ((uint64_t)qidx[i] << 32) | (uint32_t)flat[i].subject;
        idx[i] = (uint32_t)i;
    }
}

// One thread per (query, subject) group: the chain span of every region.
__global__ __launch_bounds__(kT) void chain_kernel(const uint64_t* gkeys, const uint32_t* gidx, uint64_t n,
                                                   const saber_hip_region* flat, int32_t hsp_threshold,
                                                   int64_t* span_lo, int64_t* span_hi, unsigned int* err) {
    for (uint64_t j = (uint64_t)blockIdx.x * kT + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kT) {
        if (j > 0 && gkeys[j - 1] == gkeys[j]) continue;
        uint32_t pos[kMaxGroup];
        uint32_t g = 0;
        uint64_t e = j;
        for (; e < n && gkeys[e] == gkeys[j]; ++e) {
            if (g < kMaxGroup) pos[g] = gidx[e];
            ++g;
        }
        if (g > kMaxGroup) {
            atomicOr(err, 1u);
            continue;
        }
        int32_t top = 0;
        for (uint32_t k = 0; k < g; ++k) {
            const saber_hip_region& r = flat[pos[k]];
            top = r.best > top ? r.best : top;
            span_lo[pos[k]] = r.d_min;
            span_hi[pos[k]] = r.d_max;
        }
        const int32_t cut = saber_plan::strong_cut(hsp_threshold, top);
        // strong regions in (d_min, d_max, position) order (insertion sort)
        uint32_t st[kMaxGroup];
        uint32_t ns = 0;
        for (uint32_t k = 0; k < g; ++k) {
            const uint32_t p = pos[k];
            const saber_hip_region& r = flat[p];
            if (r.best < cut) continue;
            uint32_t x = ns++;
            while (x > 0) {
                const saber_hip_region& o = flat[st[x - 1]];
                const bool after = o.d_min != r.d_min ? o.d_min > r.d_min
                                                      : (o.d_max != r.d_max ? o.d_max > r.d_max : st[x - 1] > p);
                if (!after) break;
                st[x] = st[x - 1];
                --x;
            }
            st[x] = p;
        }
        for (uint32_t s0 = 0; s0 < ns;) {
            const int64_t lo = flat[st[s0]].d_min;
            int64_t hi = flat[st[s0]].d_max;
            uint32_t s1 = s0 + 1;
            while (s1 < ns && flat[st[s1]].d_min - hi <= saber_plan::kBandBridge) {
                hi = flat[st[s1]].d_max > hi ? flat[st[s1]].d_max : hi;
                ++s1;
            }
            for (uint32_t k = s0; k < s1; ++k) {
                span_lo[st[k]] = lo;
                span_hi[st[k]] = hi;
            }
            s0 = s1;
        }
    }
}

struct PlanArgs {
    const saber_hip_region* flat;
    const uint32_t* qidx;
    const int64_t* span_lo;
    const int64_t* span_hi;
    uint64_t n;
    const uint64_t* qlen;
    const uint64_t* soff;
    saber_hip_plan_params pp;
    const saber_hip_query_stats* qs;
    const int64_t* region_floor;
    double region_e, frac;
    saber_hip_plan* plans;
    uint64_t* slen;
    int64_t* cut;
    uint8_t* amb;
    unsigned long long* nbanded;
    unsigned long long* namb;
};

// Each region's plan, subject length and two-stage cut (the smallest score at
// or above score_prefilter_frac x floor_lower_bound).
__global__ __launch_bounds__(kT) void plan_kernel(PlanArgs a) {
    unsigned long long nb = 0, na = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kT) {
        const saber_hip_region& r = a.flat[i];
        const uint32_t q = a.qidx[i];
        const uint64_t sl = a.soff[r.subject + 1] - a.soff[r.subject];
        const int64_t m = (int64_t)a.qlen[q];
        const saber_hip_plan p =
            saber_plan::plan_region(r.d_min, r.d_max, r.best, a.span_lo[i], a.span_hi[i], m, (int64_t)sl, a.pp);
        a.plans[i] = p;
        a.slen[i] = sl;
        nb += p.banded ? 1 : 0;
        const saber_hip_query_stats& qs = a.qs[q];
        bool amb = false;
        double f;
        if (!qs.per_subject) {
            f = (double)a.region_floor[q];
        } else {
            f = 1.0;
            if (a.region_e > 0.0 && isfinite(a.region_e)) {
                const double v = log(qs.k * space_lb(qs, sl, &amb) / a.region_e) / qs.lambda;
                f = v > f ? v : f;
            }
            if (qs.even_scores_only) f -= 1.0;
        }
        const double thr = a.frac * f;
        if (qs.per_subject && near_integer(thr)) amb = true;
        const double c = ceil(thr);
        a.cut[i] = c > 4.0e9 ? (int64_t)4000000000ll : (c < -4.0e9 ? (int64_t)-4000000000ll : (int64_t)c);
        a.amb[i] = amb ? 1 : 0;
        na += amb ? 1 : 0;
    }
    nb = wave_sum(nb);
    na = wave_sum(na);
    if (__lane_id() == 0) {
        if (nb) atomicAdd(a.nbanded, nb);
        if (na) atomicAdd(a.namb, na);
    }
}

__global__ __launch_bounds__(kT) void set_cut_kernel(const uint32_t* pos, const int64_t* val, uint64_t n,
                                                     int64_t* cut) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT)
        cut[pos[i]] = val[i];
}

// ---- prefilter ----

struct FeatRecArgs {
    const saber_hip_region* flat;
    const saber_hip_plan* plans;
    const uint32_t* qidx;
    const uint64_t* qlen;
    const uint64_t* soff;
    const uint64_t* slen;
    const float* rthr;
    uint64_t n;
    SearchFeatRegion* out;
};

__global__ __launch_bounds__(kT) void feature_records_kernel(FeatRecArgs a) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kT) {
        const saber_hip_region& r = a.flat[i];
        const saber_hip_plan& pl = a.plans[i];
        const uint32_t q = a.qidx[i];
        SearchFeatRegion x;
        x.s_off = a.soff[r.subject] + pl.lo;
        x.m = (uint32_t)a.qlen[q];
        x.n = (uint32_t)(pl.hi - pl.lo);
        x.band_lo = pl.banded ? (int32_t)pl.band_lo : -(int32_t)x.m + 1;
        x.band_hi = pl.banded ? (int32_t)pl.band_hi : (int32_t)x.n - 1;
        x.best = r.best;
        x.hsp_sum = r.hsp_sum;
        x.n_hsp = r.n_hsp;
        x.query = q;
        x.d_span = (int32_t)(r.d_max - r.d_min);
        x.thr = a.rthr[q];
        x.slen = (uint32_t)a.slen[i];
        x.banded = pl.banded ? 1u : 0u;
        a.out[i] = x;
    }
}

__global__ __launch_bounds__(kT) void prune_kernel(const float* logits, uint64_t n, float threshold, uint8_t* pruned,
                                                   unsigned long long* npruned) {
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        const uint8_t p = logits[i] < threshold ? 1 : 0;
        pruned[i] = p;
        c += p;
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd(npruned, c);
}

// ---- scoring passes ----

// kind 0: not scored (pruned); 1: narrow pass; 2: full pass, banded plan; 3: full pass, whole rectangle
__global__ __launch_bounds__(kT) void kind_kernel(const saber_hip_plan* plans, const uint8_t* pruned, uint64_t n,
                                                  int two_stage, uint8_t* narrow, uint8_t* full) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        int k = 0;
        if (!(pruned && pruned[i])) k = !plans[i].banded ? 3 : (two_stage ? 1 : 2);
        narrow[i] = k == 1 ? 1 : 0;
        full[i] = k >= 2 ? 1 : 0;
    }
}

struct JobArgs {
    const uint32_t* of;       // region index of each job
    uint64_t n;
    const saber_hip_region* flat;
    const saber_hip_plan* plans;
    const uint32_t* qidx;
    const uint64_t* qlen;
    int64_t narrow_band;      // > 0: narrow-pass band around the region's diagonals
    saber_hip_band_job* out;
};

__global__ __launch_bounds__(kT) void jobs_kernel(JobArgs a) {
    for (uint64_t k = (uint64_t)blockIdx.x * kT + threadIdx.x; k < a.n; k += (uint64_t)gridDim.x * kT) {
        const uint32_t i = a.of[k];
        const saber_hip_region& r = a.flat[i];
        const saber_hip_plan& p = a.plans[i];
        const int64_t m = (int64_t)a.qlen[a.qidx[i]];
        const int64_t nw = (int64_t)(p.hi - p.lo);
        saber_hip_band_job j;
        j.query = a.qidx[i];
        j.subject = (uint32_t)r.subject;
        j.win_lo = (uint32_t)p.lo;
        j.win_hi = (uint32_t)p.hi;
        if (a.narrow_band > 0) {
            j.band_lo = (int32_t)saber_plan::imax(r.d_min - a.narrow_band - (int64_t)p.lo, -(m - 1));
            j.band_hi = (int32_t)saber_plan::imin(r.d_max + a.narrow_band - (int64_t)p.lo, nw - 1);
        } else if (p.banded) {
            j.band_lo = (int32_t)p.band_lo;
            j.band_hi = (int32_t)p.band_hi;
        } else {
            j.band_lo = (int32_t)(-(m - 1));
            j.band_hi = (int32_t)(nw - 1);
        }
        a.out[k] = j;
    }
}

// The narrow pass's cut: at or above it the region is rescored over its
// plan's band, else the narrow score stands. Sums the pass's band cells.
__global__ __launch_bounds__(kT) void narrow_cut_kernel(const uint32_t* narrow_of, uint64_t n, const int32_t* score,
                                                        const uint64_t* cells, const int64_t* cut, int32_t* out,
                                                        uint8_t* resc_job, uint8_t* rescored,
                                                        unsigned long long* ncells) {
    unsigned long long cs = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kT) {
        const uint32_t i = narrow_of[k];
        const int32_t s = score[k];
        const bool again = (int64_t)s >= cut[i];
        resc_job[k] = again ? 1 : 0;
        if (again)
            rescored[i] = 1;
        else
            out[i] = s;
        cs += cells[k];
    }
    cs = wave_sum(cs);
    if (__lane_id() == 0 && cs) atomicAdd(ncells, cs);
}

__global__ __launch_bounds__(kT) void gather_u32_kernel(const uint32_t* src, const uint32_t* pos, uint64_t n,
                                                        uint32_t* dst) {
    for (uint64_t k = (uint64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kT) dst[k] = src[pos[k]];
}

__global__ __launch_bounds__(kT) void full_scatter_kernel(const uint32_t* full_of, uint64_t n, const int32_t* score,
                                                          const uint64_t* cells, const saber_hip_plan* plans,
                                                          int32_t* out, unsigned long long* ncells) {
    unsigned long long cs = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kT) {
        const uint32_t i = full_of[k];
        out[i] = score[k];
        if (plans[i].banded) cs += cells[k];
    }
    cs = wave_sum(cs);
    if (__lane_id() == 0 && cs) atomicAdd(ncells, cs);
}

// ---- candidates ----

struct CandArgs {
    const int32_t* score;
    const uint32_t* qidx;
    const uint64_t* slen;
    const saber_hip_query_stats* qs;
    const int64_t* floor;
    double ecut;
    int per_subject;
    uint64_t n;
    uint8_t* flag;
};

// A region the host's reporting loop could take (engine.rs:1138-1161): in
// the per-subject model a positive score whose E-value lower bound is within
// 1e-6 of the cut or below (every region with an ambiguous bound included),
// in the database model a score at or above the query's floor.
__global__ __launch_bounds__(kT) void cand_kernel(CandArgs a) {
    for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * kT) {
        const int64_t s = a.score[i];
        const uint32_t q = a.qidx[i];
        bool keep;
        if (!a.per_subject) {
            keep = s >= a.floor[q];
        } else if (s <= 0) {
            keep = false;
        } else {
            const saber_hip_query_stats& qs = a.qs[q];
            bool amb = false;
            const double sp = space_lb(qs, a.slen[i], &amb);
            const int64_t adj = (qs.even_scores_only && (s % 2) != 0) ? s - 1 : s;
            const double e = qs.k * sp * exp(-qs.lambda * (double)adj);
            keep = amb || !(e > a.ecut * (1.0 + 1e-6));
        }
        a.flag[i] = keep ? 1 : 0;
    }
}

__global__ __launch_bounds__(kT) void cand_gather_kernel(const uint32_t* pos, uint64_t n, const saber_hip_region* flat,
                                                         const saber_hip_plan* plans, const int32_t* score,
                                                         const uint32_t* qidx, const uint64_t* slen,
                                                         SearchCand* out) {
    for (uint64_t k = (uint64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kT) {
        const uint32_t i = pos[k];
        SearchCand c;
        c.ri = i;
        c.r = flat[i];
        c.p = plans[i];
        c.score = score[i];
        c.q = qidx[i];
        c.slen = slen[i];
        out[k] = c;
    }
}

// rocprim's two-call pattern with a reusable temporary buffer.
template <class F>
int rp(DevBuf<uint8_t>& tmp, hipStream_t st, F&& f) {
    size_t bytes = 0;
    HIP_TRY(f(nullptr, bytes));
    if (bytes > tmp.n) {
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(tmp.alloc(bytes));
    }
    HIP_TRY(f(tmp.p, bytes));
    return SABER_HIP_OK;
}

// Indices i < n with flag[i] set, in order, into out; their count on the host.
int select_flags(DevBuf<uint8_t>& tmp, hipStream_t st, const uint8_t* flag, uint64_t n, uint32_t* out,
                 DevBuf<uint64_t>& dcount, uint64_t* count) {
    rocprim::counting_iterator<uint32_t> idx(0);
    if (int r = rp(tmp, st, [&](void* t, size_t& b) {
            return rocprim::select(t, b, idx, flag, out, dcount.p, (size_t)n, st);
        }))
        return r;
    HIP_TRY(hipMemcpyAsync(count, dcount.p, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

template <class T>
int download(std::vector<T>& v, const T* d, uint64_t n, hipStream_t st) {
    v.resize(n);
    if (n) HIP_TRY(hipMemcpyAsync(v.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    return SABER_HIP_OK;
}

}  // namespace

int search_score_device(saber_hip_ctx* c, const SearchScoreIn& in, SearchScoreOut* out) {
    hipStream_t st = in.st;
    const saber_hip_search_params* P = in.P;
    const uint32_t nq = in.nq;
    const uint64_t nr = in.nregions;
    DevBuf<uint8_t> tmp;
    DevBuf<uint64_t> dcount;
    DevBuf<unsigned long long> ctr;   // [0] banded plans [1] ambiguous cuts [2] pruned [3] narrow cells [4] full cells
    HIP_TRY(dcount.alloc(1));
    HIP_TRY(ctr.alloc(8));
    HIP_TRY(hipMemsetAsync(ctr.p, 0, 8 * sizeof(unsigned long long), st));
    if (nr >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "more than 2^32 regions: split the query batch");

    // ---- cap ----
    auto t0 = Clock::now();
    DevBuf<uint64_t> keys, keys2;
    DevBuf<uint32_t> idx, idx2, qcnt;
    HIP_TRY(keys.alloc(std::max<uint64_t>(nr, 1)));
    HIP_TRY(keys2.alloc(std::max<uint64_t>(nr, 1)));
    HIP_TRY(idx.alloc(std::max<uint64_t>(nr, 1)));
    HIP_TRY(idx2.alloc(std::max<uint64_t>(nr, 1)));
    HIP_TRY(qcnt.alloc(std::max<uint32_t>(nq, 1)));
    HIP_TRY(fill_u32(qcnt.p, nq, 0, st));
    if (nr) {
        SABER_LAUNCH(cap_keys_kernel, dim3(grid_of(nr)), dim3(kT), 0, st, in.regions, nr, keys.p, idx.p, qcnt.p);
        HIP_TRY(hipGetLastError());
        const unsigned end_bit = 31 + (unsigned)bits_of(nq);
        if (int r = rp(tmp, st, [&](void* t, size_t& b) {
                return rocprim::radix_sort_pairs(t, b, keys.p, keys2.p, idx.p, idx2.p, (size_t)nr, 0u, end_bit, st);
            }))
            return r;
    }
    std::vector<uint32_t> cnt(nq);
    if (nq) HIP_TRY(hipMemcpyAsync(cnt.data(), qcnt.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<uint64_t> qstart(nq + 1, 0);
    out->roff.assign(nq + 1, 0);
    for (uint32_t q = 0; q < nq; ++q) {
        qstart[q + 1] = qstart[q] + cnt[q];
        if (cnt[q] > P->max_candidates) out->dropped_by_cap += cnt[q] - P->max_candidates;
        out->roff[q + 1] = out->roff[q] + std::min<uint64_t>(cnt[q], P->max_candidates);
    }
    const uint64_t n = out->roff[nq];
    out->regions = n;
    DevBuf<uint64_t> d_qstart, d_roff;
    if (int r = upload(d_qstart, qstart.data(), qstart.size(), st)) return r;
    if (int r = upload(d_roff, out->roff.data(), out->roff.size(), st)) return r;
    DevBuf<saber_hip_region> flat;
    DevBuf<uint32_t> qidx;
    HIP_TRY(flat.alloc(std::max<uint64_t>(n, 1)));
    HIP_TRY(qidx.alloc(std::max<uint64_t>(n, 1)));
    if (nr) {
        SABER_LAUNCH(cap_scatter_kernel, dim3(grid_of(nr)), dim3(kT), 0, st, keys2.p, idx2.p, nr, in.regions,
                     d_qstart.p, d_roff.p, (uint64_t)P->max_candidates, flat.p, qidx.p);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(st));
    out->ms_regions = ms_since(t0);

    // ---- plans and thresholds ----
    t0 = Clock::now();
    DevBuf<int64_t> span_lo, span_hi, cut;
    DevBuf<saber_hip_plan> plans;
    DevBuf<uint64_t> slen;
    DevBuf<uint8_t> amb;
    DevBuf<unsigned int> err;
    DevBuf<saber_hip_query_stats> d_qs;
    DevBuf<int64_t> d_rfloor, d_floor;
    const uint64_t n1 = std::max<uint64_t>(n, 1);
    HIP_TRY(span_lo.alloc(n1));
    HIP_TRY(span_hi.alloc(n1));
    HIP_TRY(cut.alloc(n1));
    HIP_TRY(plans.alloc(n1));
    HIP_TRY(slen.alloc(n1));
    HIP_TRY(amb.alloc(n1));
    HIP_TRY(err.alloc(1));
    HIP_TRY(fill_u32(err.p, 1, 0, st));
    if (int r = upload(d_qs, in.qstats, nq, st)) return r;
    if (int r = upload(d_rfloor, in.region_floor, nq, st)) return r;
    if (int r = upload(d_floor, in.floor, nq, st)) return r;
    if (n) {
        SABER_LAUNCH(group_keys_kernel, dim3(grid_of(n)), dim3(kT), 0, st, flat.p, qidx.p, n, keys.p, idx.p);
        HIP_TRY(hipGetLastError());
        const unsigned end_bit = 32 + (unsigned)bits_of(nq);
        if (int r = rp(tmp, st, [&](void* t, size_t& b) {
                return rocprim::radix_sort_pairs(t, b, keys.p, keys2.p, idx.p, idx2.p, (size_t)n, 0u, end_bit, st);
            }))
            return r;
        SABER_LAUNCH(chain_kernel, dim3(grid_of(n)), dim3(kT), 0, st, keys2.p, idx2.p, n, flat.p, P->hsp_threshold,
                     span_lo.p, span_hi.p, err.p);
        HIP_TRY(hipGetLastError());
        PlanArgs pa{};
        pa.flat = flat.p;
        pa.qidx = qidx.p;
        pa.span_lo = span_lo.p;
        pa.span_hi = span_hi.p;
        pa.n = n;
        pa.qlen = in.d_qlen;
        pa.soff = c->db_off_dev.p;
        pa.pp = saber_hip_plan_params{P->window_slack, P->score_band, P->hsp_threshold, 0};
        pa.qs = d_qs.p;
        pa.region_floor = d_rfloor.p;
        pa.region_e = in.region_e;
        pa.frac = P->score_prefilter_frac;
        pa.plans = plans.p;
        pa.slen = slen.p;
        pa.cut = cut.p;
        pa.amb = amb.p;
        pa.nbanded = ctr.p + 0;
        pa.namb = ctr.p + 1;
        SABER_LAUNCH(plan_kernel, dim3(grid_of(n)), dim3(kT), 0, st, pa);
        HIP_TRY(hipGetLastError());
    }
    unsigned long long hc[8];
    unsigned int herr = 0;
    HIP_TRY(hipMemcpyAsync(hc, ctr.p, sizeof hc, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&herr, err.p, sizeof herr, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (herr) return fail(SABER_HIP_EDEVICE, "more regions of one subject than max_regions_per_subject");
    out->band_scored = hc[0];
    if (hc[1]) {
        // thresholds too close to an integer for the device's log: the host's
        DevBuf<uint32_t> apos;
        HIP_TRY(apos.alloc(n1));
        uint64_t na = 0;
        if (int r = select_flags(tmp, st, amb.p, n, apos.p, dcount, &na)) return r;
        std::vector<uint32_t> pos, q;
        std::vector<uint64_t> sl;
        if (int r = download(pos, apos.p, na, st)) return r;
        HIP_TRY(hipStreamSynchronize(st));
        q.resize(na);
        sl.resize(na);
        for (uint64_t k = 0; k < na; ++k) {
            HIP_TRY(hipMemcpyAsync(&q[k], qidx.p + pos[k], sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(&sl[k], slen.p + pos[k], sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        std::vector<int64_t> val(na);
        for (uint64_t k = 0; k < na; ++k) {
            const double thr = P->score_prefilter_frac *
                               saber_hip_stats_floor_lower_bound(&in.qstats[q[k]], in.region_e, NAN, sl[k],
                                                                 in.region_floor[q[k]]);
            val[k] = (int64_t)std::ceil(thr);
        }
        DevBuf<int64_t> dval;
        if (int r = upload(dval, val.data(), na, st)) return r;
        SABER_LAUNCH(set_cut_kernel, dim3(grid_of(na)), dim3(kT), 0, st, apos.p, dval.p, na, cut.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
    }
    out->ms_plan = ms_since(t0);

    // ---- learned prefilter ----
    t0 = Clock::now();
    DevBuf<uint8_t> pruned;
    if (in.prefilter && n) {
        if (mlp_context(in.prefilter->model) != c) return fail(SABER_HIP_EINVAL, "prefilter model of another context");
        DevBuf<SearchFeatRegion> rec;
        DevBuf<float> d_rthr, d_mean, d_inv, d_qcomp, logits;
        DevBuf<__bf16> X;
        std::vector<float> rthr(nq);
        for (uint32_t q = 0; q < nq; ++q) rthr[q] = (float)in.region_thr[q];
        if (int r = upload(d_rthr, rthr.data(), nq, st)) return r;
        HIP_TRY(rec.alloc(n));
        FeatRecArgs fr{flat.p, plans.p, qidx.p, in.d_qlen, c->db_off_dev.p, slen.p, d_rthr.p, n, rec.p};
        SABER_LAUNCH(feature_records_kernel, dim3(grid_of(n)), dim3(kT), 0, st, fr);
        HIP_TRY(hipGetLastError());
        const bool norm = in.prefilter->mean && in.prefilter->inv_std;
        if (norm) {
            if (int r = upload(d_mean, in.prefilter->mean, 64, st)) return r;
            if (int r = upload(d_inv, in.prefilter->inv_std, 64, st)) return r;
        }
        HIP_TRY(d_qcomp.alloc(20ull * std::max<uint32_t>(nq, 1)));
        HIP_TRY(X.alloc(64 * n));
        SearchFeatureArgs fa{};
        fa.regions = rec.p;
        fa.n = n;
        fa.sres = c->db.p + kGuard;
        fa.mean = norm ? d_mean.p : nullptr;
        fa.inv_std = norm ? d_inv.p : nullptr;
        fa.out = X.p;
        fa.qcomp = d_qcomp.p;
        HIP_TRY(launch_search_features(fa, in.d_qoff, in.d_qres, nq, st));
        rec.release();
        HIP_TRY(logits.alloc(n));
        if (int r = mlp_forward_device(in.prefilter->model, X.p, n, logits.p, st)) return r;
        HIP_TRY(pruned.alloc(n));
        SABER_LAUNCH(prune_kernel, dim3(grid_of(n)), dim3(kT), 0, st, logits.p, n, in.prefilter->threshold, pruned.p,
                     ctr.p + 2);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
    }
    out->ms_prefilter = ms_since(t0);

    // ---- scores (score_windows) ----
    t0 = Clock::now();
    DevBuf<int32_t> score;
    DevBuf<uint8_t> rescored;
    HIP_TRY(score.alloc(n1));
    HIP_TRY(rescored.alloc(n1));
    HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(score.p), n, 0, st));
    HIP_TRY(hipMemsetAsync(rescored.p, 0, n1, st));
    const int64_t narrow_band = P->score_prefilter_band, full_band = P->score_band;
    const bool two_stage = narrow_band > 0 && narrow_band < full_band;
    DevBandInput dev{nullptr, in.d_qoff, nq, in.d_qres, in.qtotal, in.d_qbias};
    if (n) {
        DevBuf<uint8_t> fnarrow, ffull;
        DevBuf<uint32_t> narrow_of, full_of;
        HIP_TRY(fnarrow.alloc(n));
        HIP_TRY(ffull.alloc(n));
        HIP_TRY(narrow_of.alloc(n));
        HIP_TRY(full_of.alloc(n));
        SABER_LAUNCH(kind_kernel, dim3(grid_of(n)), dim3(kT), 0, st, plans.p, in.prefilter ? pruned.p : nullptr, n,
                     two_stage ? 1 : 0, fnarrow.p, ffull.p);
        HIP_TRY(hipGetLastError());
        uint64_t nn = 0, nf0 = 0;
        if (int r = select_flags(tmp, st, fnarrow.p, n, narrow_of.p, dcount, &nn)) return r;
        if (int r = select_flags(tmp, st, ffull.p, n, full_of.p, dcount, &nf0)) return r;
        fnarrow.release();
        ffull.release();
        // full-pass jobs: the full-band plans first, the narrow pass's rescored after
        DevBuf<saber_hip_band_job> fjobs;
        HIP_TRY(fjobs.alloc(std::max<uint64_t>(nf0 + nn, 1)));
        if (nf0) {
            JobArgs ja{full_of.p, nf0, flat.p, plans.p, qidx.p, in.d_qlen, 0, fjobs.p};
            SABER_LAUNCH(jobs_kernel, dim3(grid_of(nf0)), dim3(kT), 0, st, ja);
            HIP_TRY(hipGetLastError());
        }
        uint64_t nfull = nf0;
        if (nn) {
            DevBuf<saber_hip_band_job> njobs;
            HIP_TRY(njobs.alloc(nn));
            JobArgs ja{narrow_of.p, nn, flat.p, plans.p, qidx.p, in.d_qlen, narrow_band, njobs.p};
            SABER_LAUNCH(jobs_kernel, dim3(grid_of(nn)), dim3(kT), 0, st, ja);
            HIP_TRY(hipGetLastError());
            saber_hip_band_batch* b = nullptr;
            dev.jobs = njobs.p;
            if (int r = band_batch_create_impl(c, nullptr, nullptr, njobs.p, nn, &dev, in.scoring, 0, &b)) return r;
            std::unique_ptr<saber_hip_band_batch, void (*)(saber_hip_band_batch*)> hold(b, saber_hip_band_batch_destroy);
            if (int r = band_batch_enqueue(b)) return r;
            DevBuf<uint8_t> resc_job;
            DevBuf<uint32_t> resc_k;
            HIP_TRY(resc_job.alloc(nn));
            HIP_TRY(resc_k.alloc(nn));
            SABER_LAUNCH(narrow_cut_kernel, dim3(grid_of(nn)), dim3(kT), 0, st, narrow_of.p, nn, band_batch_scores(b),
                         band_batch_cells(b), cut.p, score.p, resc_job.p, rescored.p, ctr.p + 3);
            HIP_TRY(hipGetLastError());
            uint64_t nres = 0;
            if (int r = select_flags(tmp, st, resc_job.p, nn, resc_k.p, dcount, &nres)) return r;
            out->rescored = nres;
            if (nres) {
                // region of each rescored narrow job, appended to the full list
                SABER_LAUNCH(gather_u32_kernel, dim3(grid_of(nres)), dim3(kT), 0, st, narrow_of.p, resc_k.p, nres,
                             full_of.p + nf0);
                HIP_TRY(hipGetLastError());
                JobArgs jr{full_of.p + nf0, nres, flat.p, plans.p, qidx.p, in.d_qlen, 0, fjobs.p + nf0};
                SABER_LAUNCH(jobs_kernel, dim3(grid_of(nres)), dim3(kT), 0, st, jr);
                HIP_TRY(hipGetLastError());
            }
            nfull = nf0 + nres;
        }
        if (nfull) {
            saber_hip_band_batch* b = nullptr;
            dev.jobs = fjobs.p;
            if (int r = band_batch_create_impl(c, nullptr, nullptr, fjobs.p, nfull, &dev, in.scoring, 0, &b)) return r;
            std::unique_ptr<saber_hip_band_batch, void (*)(saber_hip_band_batch*)> hold(b, saber_hip_band_batch_destroy);
            if (int r = band_batch_enqueue(b)) return r;
            SABER_LAUNCH(full_scatter_kernel, dim3(grid_of(nfull)), dim3(kT), 0, st, full_of.p, nfull,
                         band_batch_scores(b), band_batch_cells(b), plans.p, score.p, ctr.p + 4);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(st));
        }
    }
    HIP_TRY(hipMemcpyAsync(hc, ctr.p, sizeof hc, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    out->pruned = hc[2];
    out->banded_cells = hc[3] + hc[4];
    out->ms_score = ms_since(t0);

    // ---- candidates for the reporting loop ----
    if (n) {
        DevBuf<uint8_t> fc;
        DevBuf<uint32_t> cpos;
        HIP_TRY(fc.alloc(n));
        HIP_TRY(cpos.alloc(n));
        CandArgs ca{score.p, qidx.p, slen.p, d_qs.p, d_floor.p, in.ecut, P->per_subject_evalues ? 1 : 0, n, fc.p};
        SABER_LAUNCH(cand_kernel, dim3(grid_of(n)), dim3(kT), 0, st, ca);
        HIP_TRY(hipGetLastError());
        uint64_t ncand = 0;
        if (int r = select_flags(tmp, st, fc.p, n, cpos.p, dcount, &ncand)) return r;
        DevBuf<SearchCand> dc;
        HIP_TRY(dc.alloc(std::max<uint64_t>(ncand, 1)));
        if (ncand) {
            SABER_LAUNCH(cand_gather_kernel, dim3(grid_of(ncand)), dim3(kT), 0, st, cpos.p, ncand, flat.p, plans.p,
                         score.p, qidx.p, slen.p, dc.p);
            HIP_TRY(hipGetLastError());
        }
        if (int r = download(out->cands, dc.p, ncand, st)) return r;
        if (in.keep_regions) {
            if (int r = download(out->flat, flat.p, n, st)) return r;
            if (int r = download(out->plans, plans.p, n, st)) return r;
            if (int r = download(out->scores, score.p, n, st)) return r;
            if (int r = download(out->rescored_flag, rescored.p, n, st)) return r;
            if (int r = download(out->qidx, qidx.p, n, st)) return r;
            if (int r = download(out->slen, slen.p, n, st)) return r;
            if (in.prefilter) {
                if (int r = download(out->pruned_flag, pruned.p, n, st)) return r;
            } else {
                out->pruned_flag.assign(n, 0);
            }
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    return SABER_HIP_OK;
}

}  // namespace saber
