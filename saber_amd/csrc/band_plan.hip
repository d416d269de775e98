// Band batch planning on the device.
//
// The plan is the host planner's (saber_hip_band_batch_create, api.hip):
// per-job geometry, transposition and shape; throughput and latency buckets;
// jobs of one query that start on the same query row paired among themselves
// (the shared pair-table flavour), the rest paired by rows; chunks of 64
// lanes dispatched in decreasing order of their estimated time. Every step
// is a per-job kernel or one of rocprim's stable radix sorts, scans and
// selections, so that a batch of millions of jobs is planned in a few
// milliseconds of device time instead of by one host thread; the host reads
// back two small summaries (totals, then the segment sizes).

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>

#include "band_plan.h"

namespace saber {

namespace {

// (bucket, flavour) keys: buckets 2 * shape (throughput) and 2 * shape + 1
// (many-row jobs), each with its shared pairs (key 2 * bucket) and the rest
// (key 2 * bucket + 1).
constexpr int kMaxSegKeys = 4 * SABER_HIP_MAX_SHAPES;
constexpr uint32_t kThreads = 256;
constexpr unsigned long long kNoErr = ~0ull;

// Totals of the per-job pass, reduced on the device and read back once.
struct PlanTotals {
    unsigned long long first_err;  // min over bad jobs of (job << 2 | kind); kNoErr: none
    unsigned long long est;        // sum over packed jobs of (rows + L + 3) * L * row_cost
    unsigned long long max_qoff;   // over packed jobs (after transposition)
    unsigned long long gen_work;   // sum over int32-kernel jobs of 4 * (w + 1)
    unsigned int max_ifirst;
    unsigned int wmax;
    unsigned int npacked, ngeneric;
    unsigned int seg_count[kMaxSegKeys];
};

// One (bucket, flavour) segment of the pair list.
struct SegInfo {
    uint32_t elem_base;   // first position in the sorted job order
    uint32_t count;       // jobs
    uint32_t pair_base;   // first pair
    uint32_t npairs;
    uint32_t chunk_base;  // first chunk (before the longest-first order)
    uint32_t nchunks;
    uint32_t per_chunk;   // pairs per chunk (64 / L)
    uint32_t shape;
    uint32_t flags;
    uint32_t L;
    uint32_t row_cost;
};

__device__ inline unsigned long long umin64(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
__device__ inline unsigned int umax32(unsigned int a, unsigned int b) { return a > b ? a : b; }

__global__ void init_totals_kernel(PlanTotals* t) {
    if (threadIdx.x == 0) {
        t->first_err = kNoErr;
        t->est = t->max_qoff = t->gen_work = 0;
        t->max_ifirst = t->wmax = t->npacked = t->ngeneric = 0;
    }
    for (int k = threadIdx.x; k < kMaxSegKeys; k += blockDim.x) t->seg_count[k] = 0;
}

struct JobArgs {
    const saber_hip_band_job* jobs;
    uint64_t njobs;
    const uint64_t* qoffs;
    uint32_t qcount;
    const uint64_t* soffs;
    uint32_t nsub;
    const int32_t* choice;
    uint32_t nchoice;
    const PlanShape* shapes;
    uint32_t packed_ok, swap_ok;
    BandJobDev* out;
    uint64_t* cells;
    int8_t* shape_of;
    uint8_t* pk_flag;   // 1: packed kernel
    uint8_t* gen_flag;  // 1: int32 kernel
    PlanTotals* tot;
};

// Per job: validation, clipped band, rows, cells, transposition, shape
// (the host planner's job loop), and the totals.
__global__ __launch_bounds__(kThreads) void plan_jobs_kernel(JobArgs a) {
    __shared__ PlanTotals bt;   // the workgroup's totals (seg_count unused)
    if (threadIdx.x == 0) {
        bt.first_err = kNoErr;
        bt.est = bt.max_qoff = bt.gen_work = 0;
        bt.max_ifirst = bt.wmax = bt.npacked = bt.ngeneric = 0;
    }
    __syncthreads();
    unsigned long long err = kNoErr, est = 0, qmax = 0, gw = 0;
    unsigned int imax = 0, wmax = 0, npk = 0, ngen = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.njobs; k += stride) {
        const saber_hip_band_job J = a.jobs[k];
        BandJobDev d{};
        uint64_t cells = 0;
        int shape = -1;
        uint8_t pk = 0, gen = 0;
        unsigned kind = 0;
        if (J.query >= a.qcount) {
            kind = 1;
        } else if (J.subject >= a.nsub) {
            kind = 2;
        } else {
            const uint64_t slen = a.soffs[J.subject + 1] - a.soffs[J.subject];
            if (J.win_hi < J.win_lo || J.win_hi > slen) kind = 3;
        }
        if (kind) {
            err = umin64(err, ((unsigned long long)k << 2) | kind);
        } else {
            const int64_t m = (int64_t)(a.qoffs[J.query + 1] - a.qoffs[J.query]);
            const int64_t n = (int64_t)J.win_hi - J.win_lo;
            d.q_off = a.qoffs[J.query];
            d.s_off = a.soffs[J.subject] + J.win_lo;
            d.m = (uint32_t)m;
            d.n = (uint32_t)n;
            if (m > 0 && n > 0 && J.band_hi >= J.band_lo) {
                const int64_t lo = J.band_lo > -(m - 1) ? (int64_t)J.band_lo : -(m - 1);
                const int64_t hi = J.band_hi < n - 1 ? (int64_t)J.band_hi : n - 1;
                const int64_t w = hi - lo + 1;
                const int64_t i_first = hi < 0 ? -hi : 0;
                const int64_t i_last = m - 1 < n - 1 - lo ? m - 1 : n - 1 - lo;
                if (w > 0 && i_last >= i_first) {
                    d.lo = (int32_t)lo;
                    d.w = (int32_t)w;
                    d.i_first = (int32_t)i_first;
                    d.rows = (int32_t)(i_last - i_first + 1);
                    cells = band_cells(d);
                    if (a.swap_ok) {
                        // transposed: rows over the window, band [-hi, -lo]
                        const int64_t t_first = lo > 0 ? lo : 0;
                        const int64_t t_last = n - 1 < m - 1 + hi ? n - 1 : m - 1 + hi;
                        if (t_last >= t_first && t_last - t_first + 1 < d.rows) {
                            BandJobDev t{};
                            t.q_off = d.s_off;
                            t.s_off = d.q_off;
                            t.m = (uint32_t)n;
                            t.n = (uint32_t)m;
                            t.swapped = 1;
                            t.lo = (int32_t)-hi;
                            t.w = (int32_t)w;
                            t.i_first = (int32_t)t_first;
                            t.rows = (int32_t)(t_last - t_first + 1);
                            d = t;
                        }
                    }
                    const bool fits = d.rows <= (int32_t)kMaxBandRows && m < kMaxPackedLen && n < kMaxPackedLen;
                    shape = (a.packed_ok && fits && w < (int64_t)a.nchoice) ? a.choice[w] : -1;
                    if (shape < 0) {
                        gen = 1;
                        ++ngen;
                        gw += 4ull * ((unsigned long long)w + 1);
                    } else {
                        pk = 1;
                        ++npk;
                        const PlanShape s = a.shapes[shape];
                        est += (unsigned long long)(d.rows + s.L + 3) * (unsigned long long)s.L * s.row_cost;
                        qmax = umax64(qmax, d.q_off);
                        imax = umax32(imax, (unsigned)d.i_first);
                        wmax = umax32(wmax, (unsigned)d.w);
                    }
                }
            }
        }
        a.out[k] = d;
        a.cells[k] = cells;
        a.shape_of[k] = (int8_t)shape;
        a.pk_flag[k] = pk;
        a.gen_flag[k] = gen;
    }
    for (int o = 32; o > 0; o >>= 1) {
        err = umin64(err, __shfl_xor(err, o));
        est += __shfl_xor(est, o);
        qmax = umax64(qmax, __shfl_xor(qmax, o));
        gw += __shfl_xor(gw, o);
        imax = umax32(imax, __shfl_xor(imax, o));
        wmax = umax32(wmax, __shfl_xor(wmax, o));
        npk += __shfl_xor(npk, o);
        ngen += __shfl_xor(ngen, o);
    }
    // the waves' totals into the workgroup's, then one global update per
    // workgroup and field (updates of one address serialise)
    if (__lane_id() == 0) {
        if (err != kNoErr) atomicMin(&bt.first_err, err);
        if (est) atomicAdd(&bt.est, est);
        if (qmax) atomicMax(&bt.max_qoff, qmax);
        if (gw) atomicAdd(&bt.gen_work, gw);
        if (imax) atomicMax(&bt.max_ifirst, imax);
        if (wmax) atomicMax(&bt.wmax, wmax);
        if (npk) atomicAdd(&bt.npacked, npk);
        if (ngen) atomicAdd(&bt.ngeneric, ngen);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PlanTotals* t = a.tot;
        if (bt.first_err != kNoErr) atomicMin(&t->first_err, bt.first_err);
        if (bt.est) atomicAdd(&t->est, bt.est);
        if (bt.max_qoff) atomicMax(&t->max_qoff, bt.max_qoff);
        if (bt.gen_work) atomicAdd(&t->gen_work, bt.gen_work);
        if (bt.max_ifirst) atomicMax(&t->max_ifirst, bt.max_ifirst);
        if (bt.wmax) atomicMax(&t->wmax, bt.wmax);
        if (bt.npacked) atomicAdd(&t->npacked, bt.npacked);
        if (bt.ngeneric) atomicAdd(&t->ngeneric, bt.ngeneric);
    }
}

struct BucketArgs {
    const uint32_t* ids;  // packed jobs
    uint32_t n;
    const BandJobDev* jobs;
    const int8_t* shape_of;
    const int32_t* choice_lat;
    const PlanShape* shapes;
    int latency_rows;     // < 0: lat_limit rule; 0: off; else a row count
    double lat_limit;     // lane instructions above which a job is many-row
    uint8_t* bucket_of;
};

// Throughput bucket 2 * shape, or many-row bucket 2 * latency shape + 1.
__global__ __launch_bounds__(kThreads) void plan_bucket_kernel(BucketArgs a) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const uint32_t id = a.ids[i];
        const int shape = a.shape_of[id];
        const PlanShape s = a.shapes[shape];
        const int32_t rows = a.jobs[id].rows;
        bool lat;
        if (a.latency_rows < 0)
            lat = (double)((unsigned long long)(rows + s.L + 3) * s.row_cost) > a.lat_limit;
        else
            lat = a.latency_rows > 0 && rows > a.latency_rows;
        const int ls = a.choice_lat[a.jobs[id].w];
        a.bucket_of[id] = (uint8_t)(lat && ls >= 0 ? 2 * ls + 1 : 2 * shape);
    }
}

// Sort keys of the grouping order (bucket, query start, first row, rows
// descending): mode 0 all in one key; modes 1-3 the fields of a three-pass
// stable sort when they do not fit 64 bits.
struct KeyArgs {
    const uint32_t* ids;
    uint32_t n;
    const BandJobDev* jobs;
    const uint8_t* bucket_of;
    int mode;
    int bq, bi;  // bits of the query start and first-row fields
    uint64_t* keys;
};

__global__ __launch_bounds__(kThreads) void plan_key_kernel(KeyArgs a) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const uint32_t id = a.ids[i];
        const BandJobDev& d = a.jobs[id];
        const uint64_t rdesc = 65535u - (uint64_t)d.rows;
        uint64_t key;
        if (a.mode == 0)
            key = ((uint64_t)a.bucket_of[id] << (a.bq + a.bi + 16)) | (d.q_off << (a.bi + 16)) |
                  ((uint64_t)d.i_first << 16) | rdesc;
        else if (a.mode == 1)
            key = ((uint64_t)d.i_first << 16) | rdesc;
        else if (a.mode == 2)
            key = d.q_off;
        else
            key = a.bucket_of[id];
        a.keys[i] = key;
    }
}

// Run starts in the grouping order: a run is a maximal sequence of
// untransposed jobs of one bucket with the same query start and first row.
__global__ __launch_bounds__(kThreads) void plan_runs_kernel(const uint32_t* ids, uint32_t n, const BandJobDev* jobs,
                                                             const uint8_t* bucket_of, uint32_t* start) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t s = 1;
        if (i > 0) {
            const uint32_t id = ids[i], p = ids[i - 1];
            const BandJobDev &x = jobs[id], &y = jobs[p];
            s = (bucket_of[id] != bucket_of[p] || x.swapped || y.swapped || x.q_off != y.q_off ||
                 x.i_first != y.i_first)
                    ? 1u
                    : 0u;
        }
        start[i] = s;
    }
}

// first[r] = position of run r + 1's first job (runs numbered from 1 by the
// inclusive scan of the starts); first[last run] = n.
__global__ __launch_bounds__(kThreads) void plan_run_first_kernel(const uint32_t* start, const uint32_t* run, uint32_t n,
                                                                  uint32_t* first) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (start[i]) first[run[i] - 1] = i;
        if (i == n - 1) first[run[i]] = n;
    }
}

// The even part of every run is shared pairs (consecutive jobs), an odd
// last job and every transposed job go to the rest. Key: (bucket, flavour)
// then the rows of the pair's first job descending (the longer one: runs
// are in rows-descending order), so a stable sort keeps pairs together.
__global__ __launch_bounds__(kThreads) void plan_classify_kernel(const uint32_t* ids, uint32_t n, const BandJobDev* jobs,
                                                                 const uint8_t* bucket_of, const uint32_t* run,
                                                                 const uint32_t* first, int nsk, uint64_t* keys,
                                                                 PlanTotals* tot) {
    __shared__ unsigned int h[kMaxSegKeys];
    for (int k = threadIdx.x; k < kMaxSegKeys; k += blockDim.x) h[k] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t id = ids[i];
        const uint32_t r = run[i] - 1;
        const uint32_t rank = i - first[r], len = first[r + 1] - first[r];
        const bool shared = !jobs[id].swapped && rank < (len & ~1u);
        const uint32_t lead = shared ? ids[i - (rank & 1u)] : id;
        const uint64_t rdesc = 65535u - (uint64_t)jobs[lead].rows;
        const uint32_t sk = 2u * bucket_of[id] + (shared ? 0u : 1u);
        keys[i] = ((uint64_t)sk << 16) | rdesc;
        atomicAdd(&h[sk], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nsk; k += blockDim.x)
        if (h[k]) atomicAdd(&tot->seg_count[k], h[k]);
}

// Pair slots: position t of a segment goes to pair pair_base + t / 2.
__global__ __launch_bounds__(kThreads) void plan_pairs_kernel(const uint32_t* ids, const uint64_t* keys, uint32_t n,
                                                              const SegInfo* segs, uint32_t* pairs) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const SegInfo& s = segs[keys[i] >> 16];
        const uint32_t t = i - s.elem_base;
        pairs[2ull * (s.pair_base + t / 2) + (t & 1u)] = ids[i];
    }
}

// Chunk g: its segment, pairs and estimated time (row steps of its first
// pair x the shape's per-row cost); key ebound - estimate sorts longest
// first.
__global__ __launch_bounds__(kThreads) void plan_chunks_kernel(const SegInfo* segs, int nsk, uint32_t nchunks,
                                                               const uint32_t* pairs, const BandJobDev* jobs,
                                                               uint64_t ebound, BandBlock* blk, uint64_t* keys,
                                                               uint32_t* idx) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nchunks; g += gridDim.x * blockDim.x) {
        int lo = 0, hi = nsk - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (segs[mid].chunk_base + segs[mid].nchunks > g)
                hi = mid;
            else
                lo = mid + 1;
        }
        const SegInfo& s = segs[lo];
        const uint32_t p0 = (g - s.chunk_base) * s.per_chunk;
        const uint32_t cnt = s.npairs - p0 < s.per_chunk ? s.npairs - p0 : s.per_chunk;
        const uint32_t lead = pairs[2ull * (s.pair_base + p0)];
        const uint64_t steps = (uint64_t)((jobs[lead].rows + (int32_t)s.L - 1 + 3) & ~3);
        blk[g] = BandBlock{s.shape, s.pair_base + p0, cnt, s.flags};
        keys[g] = ebound - steps * s.row_cost;
        idx[g] = g;
    }
}

__global__ __launch_bounds__(kThreads) void plan_gather_blocks_kernel(const BandBlock* src, const uint32_t* idx,
                                                                      uint32_t n, BandBlock* dst) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) dst[g] = src[idx[g]];
}

__global__ __launch_bounds__(kThreads) void plan_gen_size_kernel(const uint32_t* ids, uint32_t n, const BandJobDev* jobs,
                                                                 uint64_t* sz) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        sz[i] = 4ull * ((uint64_t)jobs[ids[i]].w + 1);
}

// Every kernel here loops over its range with the grid's stride; 2048
// workgroups fill the device once.
uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kThreads - 1) / kThreads, 2048)); }

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

}  // namespace

struct BandPlanScratch {
    DevBuf<saber_hip_band_job> in;
    DevBuf<uint64_t> qoffs, soffs;
    DevBuf<int32_t> choice, choice_lat;
    DevBuf<PlanShape> shapes;
    DevBuf<int8_t> shape_of;
    DevBuf<uint8_t> bucket_of, pk_flag, gen_flag;
    DevBuf<uint32_t> ids_a, ids_b;       // packed job ids (sort ping-pong); chunk indices
    DevBuf<uint64_t> keys_a, keys_b;     // their sort keys; chunk keys
    DevBuf<uint32_t> start, run, first;  // run starts, run numbers, run first positions
    DevBuf<uint32_t> count;
    DevBuf<PlanTotals> tot;
    DevBuf<SegInfo> segs;
    DevBuf<BandBlock> blk;
    DevBuf<uint64_t> gen_sz;
    DevBuf<uint8_t> tmp;                 // rocprim temporary storage
};

void plan_scratch_free(BandPlanScratch* p) { delete p; }

namespace {

// Runs a rocprim call twice: the size query, then with the scratch storage
// (grown only after the work already enqueued on st, which may use it, is
// done).
template <class F>
int rocprim_call(BandPlanScratch& S, hipStream_t st, F&& f) {
    size_t bytes = 0;
    HIP_TRY(f(nullptr, bytes));
    if (bytes > S.tmp.n) HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(S.tmp.reserve(bytes ? bytes : 1));
    HIP_TRY(f(S.tmp.p, bytes));
    return SABER_HIP_OK;
}

// Stable sort of (keys, ids) ping-pong buffers by bits [0, bits): the
// sorted data ends up in *k and *v.
int sort_pairs(BandPlanScratch& S, uint64_t** k, uint64_t** k_alt, uint32_t** v, uint32_t** v_alt, uint32_t n, int bits,
               hipStream_t st) {
    if (bits <= 0 || n < 2) return SABER_HIP_OK;
    uint64_t* kin = *k;
    uint64_t* kout = *k_alt;
    uint32_t* vin = *v;
    uint32_t* vout = *v_alt;
    if (int r = rocprim_call(S, st, [&](void* t, size_t& b) {
            return rocprim::radix_sort_pairs(t, b, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, st);
        }))
        return r;
    std::swap(*k, *k_alt);
    std::swap(*v, *v_alt);
    return SABER_HIP_OK;
}

}  // namespace

int band_plan_device(saber_hip_ctx* c, hipStream_t st, const BandPlanInput& in, DevBuf<BandJobDev>& jobs,
                     DevBuf<uint64_t>& cells, DevBuf<uint32_t>& pairs, DevBuf<BandBlock>& blocks,
                     DevBuf<uint32_t>& gen_ids, DevBuf<uint64_t>& gen_off, BandPlan* out) {
    *out = BandPlan{};
    if (!c->plan) c->plan.reset(new BandPlanScratch);
    BandPlanScratch& S = *c->plan;
    const uint64_t n = in.njobs;
    const int nshape = (int)c->shapes.size();
    const int nsk = 4 * nshape;
    if (nshape > SABER_HIP_MAX_SHAPES) return fail(SABER_HIP_EINVAL, "too many band shapes");
    HIP_TRY(jobs.alloc(n ? n : 1));
    HIP_TRY(cells.alloc(n ? n : 1));
    if (n == 0) {
        HIP_TRY(pairs.alloc(1));
        HIP_TRY(blocks.alloc(1));
        return SABER_HIP_OK;
    }

    // inputs
    std::vector<PlanShape> shp(nshape);
    for (int k = 0; k < nshape; ++k) shp[k] = PlanShape{c->shapes[k].L, (uint32_t)c->shapes[k].row_cost()};
    const saber_hip_band_job* jobs_in = in.jobs_dev;
    if (!jobs_in) {
        if (int r = upload_keep(S.in, in.jobs, n, st)) return r;
        jobs_in = S.in.p;
    }
    const uint64_t* qoffs_in = in.qoffs_dev;
    if (!qoffs_in) {
        if (int r = upload_keep(S.qoffs, in.qoffs, (uint64_t)in.qcount + 1, st)) return r;
        qoffs_in = S.qoffs.p;
    }
    const uint64_t* soffs = in.soffs_dev;
    if (!soffs) {
        if (int r = upload_keep(S.soffs, in.soffs, (uint64_t)in.nsub + 1, st)) return r;
        soffs = S.soffs.p;
    }
    const std::vector<int>& choice = in.ends ? c->choice_ends : c->choice;
    const std::vector<int>& choice_lat = in.ends ? c->choice_lat_ends : c->choice_lat;
    if (int r = upload_keep(S.choice, reinterpret_cast<const int32_t*>(choice.data()), choice.size(), st)) return r;
    if (int r = upload_keep(S.choice_lat, reinterpret_cast<const int32_t*>(choice_lat.data()), choice_lat.size(), st))
        return r;
    if (int r = upload_keep(S.shapes, shp.data(), shp.size(), st)) return r;
    HIP_TRY(S.tot.reserve(1));
    HIP_TRY(S.shape_of.reserve(n));
    HIP_TRY(S.bucket_of.reserve(n));
    HIP_TRY(S.pk_flag.reserve(n));
    HIP_TRY(S.gen_flag.reserve(n));
    HIP_TRY(S.count.reserve(1));
    SABER_LAUNCH(init_totals_kernel, dim3(1), dim3(kThreads), 0, st, S.tot.p);
    HIP_TRY(hipGetLastError());

    // per-job pass and its totals
    JobArgs ja{};
    ja.jobs = jobs_in;
    ja.njobs = n;
    ja.qoffs = qoffs_in;
    ja.qcount = in.qcount;
    ja.soffs = soffs;
    ja.nsub = in.nsub;
    ja.choice = S.choice.p;
    ja.nchoice = (uint32_t)choice.size();
    ja.shapes = S.shapes.p;
    ja.packed_ok = in.packed_ok;
    ja.swap_ok = in.swap_ok;
    ja.out = jobs.p;
    ja.cells = cells.p;
    ja.shape_of = S.shape_of.p;
    ja.pk_flag = S.pk_flag.p;
    ja.gen_flag = S.gen_flag.p;
    ja.tot = S.tot.p;
    SABER_LAUNCH(plan_jobs_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, ja);
    HIP_TRY(hipGetLastError());
    PlanTotals tot{};
    HIP_TRY(hipMemcpyAsync(&tot, S.tot.p, sizeof(PlanTotals), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (tot.first_err != kNoErr) {
        switch (tot.first_err & 3u) {
            case 1: return fail(SABER_HIP_EINVAL, "job query index out of range");
            case 2: return fail(SABER_HIP_EINVAL, "job subject index out of range");
            default: return fail(SABER_HIP_EINVAL, "job window outside its subject");
        }
    }
    out->wmax = (int32_t)tot.wmax;
    out->ngeneric = tot.ngeneric;
    out->npacked = tot.npacked;
    out->gen_work = tot.gen_work;

    // int32-kernel jobs in job order, and their work offsets
    if (tot.ngeneric) {
        const uint32_t ng = tot.ngeneric;
        HIP_TRY(gen_ids.alloc(ng));
        HIP_TRY(gen_off.alloc(ng));
        HIP_TRY(S.gen_sz.reserve(ng));
        rocprim::counting_iterator<uint32_t> idx(0);
        uint32_t* gi = gen_ids.p;
        const uint8_t* gf = S.gen_flag.p;
        uint32_t* cnt = S.count.p;
        if (int r = rocprim_call(S, st, [&](void* t, size_t& b) {
                return rocprim::select(t, b, idx, gf, gi, cnt, (size_t)n, st);
            }))
            return r;
        SABER_LAUNCH(plan_gen_size_kernel, dim3(grid_for(ng)), dim3(kThreads), 0, st, gen_ids.p, ng, jobs.p,
                           S.gen_sz.p);
        HIP_TRY(hipGetLastError());
        const uint64_t* gs = S.gen_sz.p;
        uint64_t* go = gen_off.p;
        if (int r = rocprim_call(S, st, [&](void* t, size_t& b) {
                return rocprim::exclusive_scan(t, b, gs, go, (uint64_t)0, (size_t)ng, rocprim::plus<uint64_t>(), st);
            }))
            return r;
    }

    const uint32_t np = tot.npacked;
    if (np == 0) {
        HIP_TRY(pairs.alloc(1));
        HIP_TRY(blocks.alloc(1));
        return SABER_HIP_OK;
    }
    HIP_TRY(S.ids_a.reserve(np));
    HIP_TRY(S.ids_b.reserve(np));
    HIP_TRY(S.keys_a.reserve(np));
    HIP_TRY(S.keys_b.reserve(np));
    HIP_TRY(S.start.reserve(np));
    HIP_TRY(S.run.reserve(np));
    HIP_TRY(S.first.reserve((uint64_t)np + 1));
    uint32_t *ids = S.ids_a.p, *ids_alt = S.ids_b.p;
    uint64_t *keys = S.keys_a.p, *keys_alt = S.keys_b.p;
    {
        rocprim::counting_iterator<uint32_t> idx(0);
        const uint8_t* pf = S.pk_flag.p;
        uint32_t* cnt = S.count.p;
        if (int r = rocprim_call(S, st, [&](void* t, size_t& b) {
                return rocprim::select(t, b, idx, pf, ids, cnt, (size_t)n, st);
            }))
            return r;
    }

    // buckets
    BucketArgs ba{};
    ba.ids = ids;
    ba.n = np;
    ba.jobs = jobs.p;
    ba.shape_of = S.shape_of.p;
    ba.choice_lat = S.choice_lat.p;
    ba.shapes = S.shapes.p;
    ba.latency_rows = c->latency_rows;
    const double est_ms = 0.5 * (double)tot.est / kGpuLaneInstrPerMs;
    ba.lat_limit = kLatencyShare * est_ms * kLaneInstrPerMs;
    ba.bucket_of = S.bucket_of.p;
    SABER_LAUNCH(plan_bucket_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, ba);
    HIP_TRY(hipGetLastError());

    // grouping order: bucket, query start, first row, rows descending
    const int bb = bits_for((uint64_t)(2 * nshape - 1)), bq = bits_for(tot.max_qoff), bi = bits_for(tot.max_ifirst);
    KeyArgs ka{};
    ka.n = np;
    ka.jobs = jobs.p;
    ka.bucket_of = S.bucket_of.p;
    ka.bq = bq;
    ka.bi = bi;
    auto key_pass = [&](int mode, int bits) -> int {
        ka.ids = ids;
        ka.mode = mode;
        ka.keys = keys;
        SABER_LAUNCH(plan_key_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, ka);
        HIP_TRY(hipGetLastError());
        return sort_pairs(S, &keys, &keys_alt, &ids, &ids_alt, np, bits, st);
    };
    if (bb + bq + bi + 16 <= 64) {
        if (int r = key_pass(0, bb + bq + bi + 16)) return r;
    } else {
        if (int r = key_pass(1, bi + 16)) return r;
        if (int r = key_pass(2, bq)) return r;
        if (int r = key_pass(3, bb)) return r;
    }

    // runs, flavours and the (bucket, flavour, rows) order
    SABER_LAUNCH(plan_runs_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, ids, np, jobs.p, S.bucket_of.p,
                       S.start.p);
    HIP_TRY(hipGetLastError());
    {
        const uint32_t* sp = S.start.p;
        uint32_t* rp = S.run.p;
        if (int r = rocprim_call(S, st, [&](void* t, size_t& b) {
                return rocprim::inclusive_scan(t, b, sp, rp, (size_t)np, rocprim::plus<uint32_t>(), st);
            }))
            return r;
    }
    SABER_LAUNCH(plan_run_first_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, S.start.p, S.run.p, np,
                       S.first.p);
    HIP_TRY(hipGetLastError());
    SABER_LAUNCH(plan_classify_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, ids, np, jobs.p, S.bucket_of.p,
                       S.run.p, S.first.p, nsk, keys, S.tot.p);
    HIP_TRY(hipGetLastError());
    if (int r = sort_pairs(S, &keys, &keys_alt, &ids, &ids_alt, np, 16 + bits_for((uint64_t)(nsk - 1)), st)) return r;
    HIP_TRY(hipMemcpyAsync(tot.seg_count, S.tot.p->seg_count, sizeof(tot.seg_count), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));

    // segments: for each bucket its shared pairs, then the rest
    std::vector<SegInfo> segs(nsk);
    uint64_t elem = 0, pair = 0, chunk = 0, ebound = 0;
    for (int sk = 0; sk < nsk; ++sk) {
        const int shape = sk / 4;
        const Shape& sh = c->shapes[shape];
        SegInfo& s = segs[sk];
        s.elem_base = (uint32_t)elem;
        s.count = tot.seg_count[sk];
        s.pair_base = (uint32_t)pair;
        s.npairs = (s.count + 1) / 2;
        s.per_chunk = 64u / (uint32_t)sh.L;
        s.chunk_base = (uint32_t)chunk;
        s.nchunks = (s.npairs + s.per_chunk - 1) / s.per_chunk;
        s.shape = (uint32_t)shape;
        s.flags = (sk & 1) ? 0u : kChunkShared;
        s.L = (uint32_t)sh.L;
        s.row_cost = (uint32_t)sh.row_cost();
        elem += s.count;
        pair += s.npairs;
        chunk += s.nchunks;
        const uint64_t max_steps = (uint64_t)(((int64_t)kMaxBandRows + sh.L - 1 + 3) & ~3);
        ebound = std::max<uint64_t>(ebound, max_steps * s.row_cost);
    }
    if (elem != np) return fail(SABER_HIP_EDEVICE, "device band plan: segment sizes do not add up");
    if (pair >= 0xFFFFFFFFull) return fail(SABER_HIP_EINVAL, "too many job pairs for one batch");
    if (int r = upload_keep(S.segs, segs.data(), segs.size(), st)) return r;
    out->npairs = pair;
    out->nchunks = chunk;

    // pair list
    HIP_TRY(pairs.alloc(2 * pair));
    HIP_TRY(fill_u32(pairs.p, 2 * pair, kNone, st));
    SABER_LAUNCH(plan_pairs_kernel, dim3(grid_for(np)), dim3(kThreads), 0, st, ids, keys, np, S.segs.p, pairs.p);
    HIP_TRY(hipGetLastError());

    // chunks, longest estimated time first (stable)
    const uint32_t nch = (uint32_t)chunk;
    HIP_TRY(S.blk.reserve(nch));
    SABER_LAUNCH(plan_chunks_kernel, dim3(grid_for(nch)), dim3(kThreads), 0, st, S.segs.p, nsk, nch, pairs.p,
                       jobs.p, ebound, S.blk.p, keys, ids);
    HIP_TRY(hipGetLastError());
    if (int r = sort_pairs(S, &keys, &keys_alt, &ids, &ids_alt, nch, bits_for(ebound), st)) return r;
    HIP_TRY(blocks.alloc(nch));
    SABER_LAUNCH(plan_gather_blocks_kernel, dim3(grid_for(nch)), dim3(kThreads), 0, st, S.blk.p, ids, nch,
                       blocks.p);
    HIP_TRY(hipGetLastError());
    // the segment table's host copy stays alive until its upload is done
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

}  // namespace saber
