// C ABI (include/saber_hip.h): context, device-resident database, the band
// planner and the batch runner.
//
// Planner: every job's band is clipped to its rectangle (Band::clip,
// src/algorithm/banded.rs:45-47); empty jobs score 0. The remaining jobs are
// bucketed by clipped width into the packed kernel's variants (or the exact
// int32 kernel), sorted inside a bucket by row count (longest first, so the
// longest pairs start first and the two jobs of a lane have nearly the same
// length) and paired two per lane. The kernels write results by job id, so
// the order of work never changes an output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <atomic>
#include <memory>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/saber_hip.h"
#include "common.h"
#include "ctx.h"
#include "band_plan.h"
#include "host_par.h"

using namespace saber;

namespace saber {

namespace {
thread_local std::string g_err;
}  // namespace

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

const char* g_err_cstr() { return g_err.c_str(); }

namespace {

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Residue codes >= 24 all score as X (src/alphabet.rs:39-41); every kernel
// maps them so. Stored residue buffers hold the canonical code (X for any
// such byte), which lets the band kernel index its 24-row table directly.
__global__ void canon_codes_kernel(uint8_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint8_t c = p[i];
        if (c >= 24) p[i] = (uint8_t)saber::kCodeX;
    }
}

}  // namespace

hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 256;
    const uint64_t grid = (n + block - 1) / block;
    SABER_LAUNCH(fill_u32_kernel, dim3((uint32_t)grid), dim3(block), 0, s, p, n, v);
    return hipGetLastError();
}

int injected(const saber_hip_ctx* c) {
    if (c && (c->flags & SABER_HIP_INJECT_FAULTS))
        return fail(SABER_HIP_EDEVICE, "injected fault (SABER_HIP_INJECT_FAULTS)");
    return SABER_HIP_OK;
}

int check_seqset(const saber_hip_seqset* s, const char* what) {
    if (!s) return fail(SABER_HIP_EINVAL, std::string(what) + " is null");
    if (s->count > 0 && !s->offsets) return fail(SABER_HIP_EINVAL, std::string(what) + ": offsets null");
    if (s->count > 0 && s->offsets[s->count] > 0 && !s->residues)
        return fail(SABER_HIP_EINVAL, std::string(what) + ": residues null");
    for (uint32_t i = 0; i < s->count; ++i)
        if (s->offsets[i + 1] < s->offsets[i])
            return fail(SABER_HIP_EINVAL, std::string(what) + ": offsets not monotone");
    return SABER_HIP_OK;
}

int upload_codes(DevBuf<uint8_t>& d, const uint8_t* h, uint64_t n, hipStream_t s) {
    if (int r = upload_guarded(d, h, n, s)) return r;
    if (n) {
        const uint32_t block = 256;
        const uint64_t grid = (n + block - 1) / block;
        SABER_LAUNCH(canon_codes_kernel, dim3((uint32_t)grid), dim3(block), 0, s, d.p + kGuard, n);
        HIP_TRY(hipGetLastError());
    }
    return SABER_HIP_OK;
}

}  // namespace saber

struct saber_hip_band_batch {
    saber_hip_ctx* ctx = nullptr;
    uint64_t njobs = 0;
    bool has_bias = false;
    bool ends = true;          // end cells computed (else scores only)
    int32_t go = 0, ge = 0;
    DevBuf<uint8_t> q, s;
    DevBuf<int8_t> bias;
    DevBuf<uint32_t> table;   // int8 band table as words
    DevBuf<uint32_t> pair_table;
    DevBuf<uint32_t> next_chunk;
    DevBuf<int32_t> mat;
    DevBuf<BandJobDev> jobs;
    DevBuf<uint32_t> pairs;
    // multi-shape launch: blocks in dispatch order (longest first)
    DevBuf<BandBlock> blocks;
    uint64_t nchunks = 0;
    // exact int32 kernel: wide / unsuitable jobs known at plan time, plus the
    // packed kernel's saturated jobs known after it ran
    uint32_t ngeneric = 0;
    uint64_t npacked = 0;      // jobs of the packed kernel
    DevBuf<uint32_t> gen_ids;      // static int32 jobs (known at plan time), in job order
    DevBuf<uint64_t> gen_off;
    DevBuf<int32_t> gen_work;
    // saturated jobs are rescored by a fixed grid reading their count on the
    // device: flag_threads threads with flag_slot int32 of work each
    DevBuf<int32_t> flag_work;
    uint32_t flag_threads = 0, flag_slot = 0;
    // the whole step captured once and replayed
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t passes = 0;                   // passes enqueued so far
    DevBuf<uint32_t> flagged;
    DevBuf<uint32_t> nflagged;
    DevBuf<int32_t> score;
    DevBuf<uint32_t> q_end, s_end;
    DevBuf<uint64_t> cells;
    const uint8_t* sres = nullptr;  // own copy or the context database (data start)
    const uint8_t* qres = nullptr;  // own copy or the caller's device codes (data start)
    const int8_t* qbias = nullptr;  // own copy, the caller's device biases, or none
    uint64_t qlen_total = 0, slen_total = 0;
    bool ran = false;
    uint32_t last_flagged = 0;
};

namespace {

// Stable LSD radix sort of v by the low `bits` bits of key[v[i]] (11-bit
// digits): the planner's orderings over millions of jobs, without comparator
// sorts that chase job records across memory.
void radix_sort_by(std::vector<uint32_t>& v, const std::vector<uint64_t>& key, int bits) {
    if (v.size() < 2 || bits <= 0) return;
    std::vector<uint32_t> tmp(v.size());
    std::vector<uint64_t> k(v.size()), ktmp(v.size());
    for (size_t i = 0; i < v.size(); ++i) k[i] = key[v[i]];
    for (int shift = 0; shift < bits; shift += 11) {
        uint32_t cnt[2049] = {0};
        for (size_t i = 0; i < v.size(); ++i) cnt[((k[i] >> shift) & 2047) + 1]++;
        for (int d = 0; d < 2048; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < v.size(); ++i) {
            const uint32_t dst = cnt[(k[i] >> shift) & 2047]++;
            tmp[dst] = v[i];
            ktmp[dst] = k[i];
        }
        v.swap(tmp);
        k.swap(ktmp);
    }
}

// Autotuner measurement: the median of three timings and their spread
// (max - min). A candidate replaces the incumbent only when its median is
// lower by more than both spreads and by more than 1 %, so run-to-run noise
// does not flip the choice.
struct Timing {
    double median = 0.0, spread = 0.0;
};
template <class F>
int timed3(F&& once, Timing* out) {
    double t[3];
    for (int i = 0; i < 3; ++i)
        if (int r = once(&t[i])) return r;
    std::sort(t, t + 3);
    out->median = t[1];
    out->spread = t[2] - t[0];
    return SABER_HIP_OK;
}
bool clearly_faster(const Timing& cand, const Timing& inc) {
    return cand.median < inc.median * 0.99 && inc.median - cand.median > std::max(cand.spread, inc.spread);
}

int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

void default_choice(saber_hip_ctx* c) {
    int nv = 0;
    const BandVariant* v = band_variants(&nv);
    c->shapes.clear();
    for (int i = 0; i < nv; ++i) {
        if (v[i].bias) continue;
        c->shapes.push_back(Shape{v[i].L, v[i].C, v[i].K});
    }
    int wmax = 0;
    for (auto& s : c->shapes) wmax = std::max(wmax, s.W());
    for (int e = 0; e < 2; ++e) {
        std::vector<int>& ch = e ? c->choice_ends : c->choice;
        std::vector<int>& lat = e ? c->choice_lat_ends : c->choice_lat;
        ch.assign((size_t)wmax + 1, -1);
        lat.assign((size_t)wmax + 1, -1);
        for (int w = 1; w <= wmax; ++w) {
            int best = -1, fast = -1;
            for (int k = 0; k < (int)c->shapes.size(); ++k) {
                const Shape& s = c->shapes[k];
                if (!s.covers(w) || (e && !s.ends_ok())) continue;
                // throughput: least work per pair-row; latency: least per-row
                // time, and among equals the widest shape, so that the few
                // many-row jobs of all widths share one launch
                if (best < 0 || s.pair_cost() < c->shapes[best].pair_cost()) best = k;
                if (fast < 0 || s.row_cost() < c->shapes[fast].row_cost() ||
                    (s.row_cost() == c->shapes[fast].row_cost() && s.W() > c->shapes[fast].W()))
                    fast = k;
            }
            ch[w] = best;
            lat[w] = fast;
        }
    }
}

// Pair table of the packed kernels: TP[q][a][b] = S(q,a) | S(q,b) << 16 (int16
// halves) for query codes q <= 24 and subject codes a, b <= 24, where code 24
// (the sentinel column, the dead query row) scores -128 and the scores are
// clamped to int8 (the packed kernels only run when every score fits).
std::vector<uint32_t> build_pair_table(const int32_t* matrix) {
    std::vector<uint32_t> t(kPairWords, 0);
    auto s8 = [&](int q, int x) -> int16_t {
        if (q == kBandSentinel || x == kBandSentinel) return -128;
        return (int16_t)std::max(-128, std::min(127, matrix[q * 24 + x]));
    };
    for (int q = 0; q <= kBandSentinel; ++q)
        for (int x = 0; x <= kBandSentinel; ++x)
            for (int y = 0; y <= kBandSentinel; ++y)
                t[(size_t)q * 625 + x * 25 + y] = (uint32_t)(uint16_t)s8(q, x) | ((uint32_t)(uint16_t)s8(q, y) << 16);
    return t;
}

}  // namespace

extern "C" {

const char* saber_hip_version(void) { return "saber-mi355x 0.1 (gfx950)"; }

const char* saber_hip_last_error(void) { return saber::g_err_cstr(); }

int saber_hip_device_count(int* count) {
    if (!count) return fail(SABER_HIP_EINVAL, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return SABER_HIP_OK;
}

int saber_hip_create(int device, const saber_hip_opts* opts, saber_hip_ctx** out) {
    if (!out) return fail(SABER_HIP_EINVAL, "out is null");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(SABER_HIP_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(SABER_HIP_EINVAL, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    auto* c = new saber_hip_ctx();
    c->device = device;
    c->flags = opts ? opts->flags : 0;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus < 1)
        c->cus = 1;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < kStreams && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&c->work[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    for (int k = 0; k < kStreams && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        saber_hip_destroy(c);
        return fail(SABER_HIP_EDEVICE, std::string("stream/event creation: ") + hipGetErrorString(e));
    }
    if (int r = xdrop_noalloc_setup(device)) {
        saber_hip_destroy(c);
        return r;
    }
    default_choice(c);
    *out = c;
    return SABER_HIP_OK;
}

void saber_hip_destroy(saber_hip_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < kStreams; ++k)
        if (c->work[k]) (void)hipStreamSynchronize(c->work[k]);
    c->scan.reset();
    c->plan.reset();
    c->db.release();
    c->db_off_dev.release();
    for (int k = 0; k < kStreams; ++k) {
        if (c->work[k]) (void)hipStreamDestroy(c->work[k]);
        if (c->join[k]) (void)hipEventDestroy(c->join[k]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int saber_hip_upload_db(saber_hip_ctx* c, const saber_hip_seqset* subjects) {
    if (!c) return fail(SABER_HIP_EINVAL, "ctx is null");
    if (int r = check_seqset(subjects, "subjects")) return r;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t nres = subjects->count ? subjects->offsets[subjects->count] : 0;
    c->scan.reset();
    ++c->db_gen;
    if (int r = upload_codes(c->db, subjects->residues, nres, c->stream)) return r;
    if (int r = upload(c->db_off_dev, subjects->offsets, (uint64_t)subjects->count + 1, c->stream)) return r;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->db_off.assign(subjects->offsets, subjects->offsets + subjects->count + 1);
    c->db_host.resize(nres);
    for (uint64_t i = 0; i < nres; ++i) {
        const uint8_t x = subjects->residues[i];
        c->db_host[i] = x < 24 ? x : (uint8_t)kCodeX;
    }
    return SABER_HIP_OK;
}

int saber_hip_upload_index(saber_hip_ctx* c, const saber_hip_index* idx) {
    if (!c || !idx) return fail(SABER_HIP_EINVAL, "null argument");
    if (idx->num_subjects >= 0xFFFFFFFFull) return fail(SABER_HIP_EINVAL, "more than 2^32 - 1 subjects");
    const saber_hip_seqset s{idx->residues, idx->offsets, (uint32_t)idx->num_subjects, nullptr};
    return saber_hip_upload_db(c, &s);
}

int saber_hip_band_variant(int v, int* L, int* C, int* K, int* bias) {
    int nv = 0;
    const BandVariant* t = band_variants(&nv);
    if (v < 0 || v >= nv) return fail(SABER_HIP_EINVAL, "variant index out of range");
    if (L) *L = t[v].L;
    if (C) *C = t[v].C;
    if (K) *K = t[v].K;
    if (bias) *bias = t[v].bias;
    return SABER_HIP_OK;
}

int saber_hip_band_choose(saber_hip_ctx* c, int w_lo, int w_hi, int L, int C, int K) {
    if (!c) return fail(SABER_HIP_EINVAL, "ctx is null");
    std::lock_guard<std::mutex> g(c->mu);
    if (w_lo < 1 || w_hi < w_lo) return fail(SABER_HIP_EINVAL, "bad width range");
    int shape = -1;
    if (L != 0) {
        for (int k = 0; k < (int)c->shapes.size(); ++k)
            if (c->shapes[k].L == L && c->shapes[k].C == C && c->shapes[k].K == K) shape = k;
        if (shape < 0) return fail(SABER_HIP_EINVAL, "no compiled variant with that shape");
        const Shape& s = c->shapes[shape];
        if (!s.covers(w_lo) || !s.covers(w_hi))
            return fail(SABER_HIP_EINVAL, "widths outside the variant's coverage");
    }
    if ((size_t)w_hi >= c->choice.size()) {
        if (L != 0) return fail(SABER_HIP_EINVAL, "width above the widest variant");
        w_hi = (int)c->choice.size() - 1;
    }
    // end-cell batches take the choice too unless the shape's lanes are too
    // wide for their register copy of a row
    const bool ends_too = shape < 0 || c->shapes[shape].ends_ok();
    for (int w = w_lo; w <= w_hi; ++w) {
        c->choice[w] = c->choice_lat[w] = shape;
        if (ends_too) c->choice_ends[w] = c->choice_lat_ends[w] = shape;
    }
    return SABER_HIP_OK;
}

int saber_hip_band_tune(saber_hip_ctx* c, int latency_rows) {
    if (!c) return fail(SABER_HIP_EINVAL, "ctx is null");
    if (latency_rows < -1) return fail(SABER_HIP_EINVAL, "latency_rows must be >= -1");
    std::lock_guard<std::mutex> g(c->mu);
    c->latency_rows = latency_rows;
    return SABER_HIP_OK;
}

int saber_hip_band_autotune(saber_hip_ctx* c, const saber_hip_seqset* queries,
                            const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                            uint64_t njobs, const saber_hip_scheme* scheme, uint64_t target_cells,
                            saber_hip_tune_entry* report, int nreport, int* nclasses) {
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (njobs > 0 && !jobs) return fail(SABER_HIP_EINVAL, "jobs is null");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    if (nclasses) *nclasses = 0;
    if (target_cells == 0) target_cells = 2000000000ull;
    std::vector<Shape> shapes;
    std::vector<int> choice0;
    {
        std::lock_guard<std::mutex> g(c->mu);
        shapes = c->shapes;
        choice0 = c->choice;
    }
    int wmax = 0;
    for (const Shape& sh : shapes) wmax = std::max(wmax, sh.W());
    // Classes: maximal runs of widths covered by the same set of shapes.
    std::vector<std::pair<int, int>> classes;
    std::vector<int> wclass((size_t)wmax + 1, -1);
    std::vector<bool> prev;
    for (int w = 1; w <= wmax; ++w) {
        std::vector<bool> cov(shapes.size());
        bool any = false;
        for (size_t k = 0; k < shapes.size(); ++k) any |= (cov[k] = shapes[k].covers(w));
        if (!any) {
            prev.clear();
            continue;
        }
        if (cov != prev) classes.push_back({w, w});
        classes.back().second = w;
        wclass[w] = (int)classes.size() - 1;
        prev = cov;
    }
    // The sample: the batch's first jobs in order, up to target_cells band
    // cells, so that every trial times the same mix of widths, pairings and
    // queries as the batch itself.
    const uint64_t* soffs = subjects ? subjects->offsets : c->db_off.data();
    const uint32_t nsub = subjects ? subjects->count : (c->db_off.empty() ? 0 : (uint32_t)c->db_off.size() - 1);
    std::vector<uint64_t> ccells(classes.size(), 0);
    uint64_t nsample = 0, tot = 0;
    for (; nsample < njobs && tot < target_cells; ++nsample) {
        const saber_hip_band_job& J = jobs[nsample];
        if (J.query >= queries->count || J.subject >= nsub) return fail(SABER_HIP_EINVAL, "job index out of range");
        const uint64_t slen = soffs[J.subject + 1] - soffs[J.subject];
        if (J.win_hi < J.win_lo || J.win_hi > slen) return fail(SABER_HIP_EINVAL, "job window outside its subject");
        const int64_t m = (int64_t)(queries->offsets[J.query + 1] - queries->offsets[J.query]);
        const int64_t n = (int64_t)J.win_hi - J.win_lo;
        if (m == 0 || n == 0 || J.band_hi < J.band_lo) continue;
        const int64_t lo = std::max<int64_t>(J.band_lo, -(m - 1));
        const int64_t hi = std::min<int64_t>(J.band_hi, n - 1);
        const int64_t w = hi - lo + 1;
        const int64_t i_first = std::max<int64_t>(0, -hi), i_last = std::min<int64_t>(m - 1, n - 1 - lo);
        if (w <= 0 || i_last < i_first) continue;
        BandJobDev d{};
        d.m = (uint32_t)m;
        d.n = (uint32_t)n;
        d.lo = (int32_t)lo;
        d.w = (int32_t)w;
        d.i_first = (int32_t)i_first;
        d.rows = (int32_t)(i_last - i_first + 1);
        const uint64_t cells = band_cells(d);
        tot += cells;
        if (w <= wmax && wclass[w] >= 0) ccells[wclass[w]] += cells;
    }
    // Coordinate descent from the modelled choice: the classes in decreasing
    // share of the sample's cells; per class every covering shape, kept when
    // the sample's packed launch gets clearly faster (clearly_faster: median
    // of three timings lower by more than 1 % and by more than their spread).
    std::vector<int> order;
    for (size_t cl = 0; cl < classes.size(); ++cl)
        if (ccells[cl] * 100 >= tot && ccells[cl] > 0) order.push_back((int)cl);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return ccells[x] > ccells[y]; });
    int saved_rows = -1;
    std::vector<int> saved_lat;
    {
        std::lock_guard<std::mutex> g(c->mu);
        saved_rows = c->latency_rows;
        saved_lat = c->choice_lat;
    }
    auto trial = [&](Timing* out) -> int {
        saber_hip_band_batch* b = nullptr;
        if (int r = saber_hip_band_batch_create(c, queries, subjects, jobs, nsample, scheme, 0, &b)) return r;
        bool warm = false;
        const int r = timed3([&](double* ms_out) -> int {
            double ms = 0.0, band_ms = 0.0;
            const int rr = saber_hip_band_batch_time(b, warm ? 0 : 1, 3, &ms, &band_ms, nullptr, nullptr, nullptr);
            warm = true;
            *ms_out = band_ms;
            return rr;
        }, out);
        saber_hip_band_batch_destroy(b);
        return r;
    };
    Timing best;
    int status = trial(&best);
    double best_ms = best.median;
    std::vector<int> chosen(classes.size(), -1);
    for (int cl : order) {
        if (status != SABER_HIP_OK) break;
        const auto [w_lo, w_hi] = classes[cl];
        const int start = choice0[w_lo];
        chosen[cl] = start;
        for (size_t k = 0; k < shapes.size(); ++k) {
            if ((int)k == start || !shapes[k].covers(w_lo) || !shapes[k].covers(w_hi)) continue;
            {
                std::lock_guard<std::mutex> g(c->mu);
                for (int w = w_lo; w <= w_hi; ++w) c->choice[w] = (int)k;
            }
            Timing t;
            status = trial(&t);
            if (status != SABER_HIP_OK) break;
            if (clearly_faster(t, best)) {
                best = t;
                best_ms = t.median;
                chosen[cl] = (int)k;
            }
        }
        std::lock_guard<std::mutex> g(c->mu);
        for (int w = w_lo; w <= w_hi; ++w) c->choice[w] = chosen[cl] >= 0 ? chosen[cl] : choice0[w];
    }
    {
        std::lock_guard<std::mutex> g(c->mu);
        c->latency_rows = saved_rows;
        c->choice_lat = saved_lat;
        if (status != SABER_HIP_OK) c->choice = choice0;
    }
    if (status != SABER_HIP_OK) return status;
    int nout = 0;
    for (int cl : order) {
        if (chosen[cl] < 0) continue;
        if (report && nout < nreport) {
            saber_hip_tune_entry& e = report[nout];
            e.w_lo = classes[cl].first;
            e.w_hi = classes[cl].second;
            e.L = shapes[chosen[cl]].L;
            e.C = shapes[chosen[cl]].C;
            e.K = shapes[chosen[cl]].K;
            e.reserved = chosen[cl] != choice0[classes[cl].first];
            e.cells = ccells[cl];
            e.gcups = best_ms > 0.0 ? (double)tot / best_ms / 1e6 : 0.0;
        }
        ++nout;
    }
    if (nclasses) *nclasses = nout;
    return SABER_HIP_OK;
}

}  // extern "C"

namespace {

// The host planner: the plan of band_plan_device (band_plan.hip) built by
// one host thread, for small batches (and every batch of a context created
// with SABER_HIP_HOST_PLAN), then uploaded into the batch's buffers.
template <class Lap>
int plan_on_host(saber_hip_ctx* c, hipStream_t st, const BandPlanInput& in, saber_hip_band_batch* b, BandPlan* out,
                 Lap&& lap) {
    const uint64_t njobs = in.njobs;
    std::vector<BandJobDev> host_jobs(njobs);
    std::vector<BandBlock> host_blocks;
    std::vector<uint32_t> generic;  // int32-kernel jobs, in job order
    std::vector<uint64_t> host_cells(njobs, 0);
    std::vector<int> thr_shape(njobs, -1);  // throughput shape of each packed job
    double est_work = 0.0;                  // lane instructions, all packed jobs
    // bucket 2 * shape + 1 holds the shape's many-row jobs
    std::vector<std::vector<uint32_t>> bucket(2 * c->shapes.size());
    std::vector<int32_t> rows_of(njobs, 0);
    for (uint64_t k = 0; k < njobs; ++k) {
        const saber_hip_band_job& J = in.jobs[k];
        if (J.query >= in.qcount) return fail(SABER_HIP_EINVAL, "job query index out of range");
        if (J.subject >= in.nsub) return fail(SABER_HIP_EINVAL, "job subject index out of range");
        const uint64_t slen = in.soffs[J.subject + 1] - in.soffs[J.subject];
        if (J.win_hi < J.win_lo || J.win_hi > slen)
            return fail(SABER_HIP_EINVAL, "job window outside its subject");
        const int64_t m = (int64_t)(in.qoffs[J.query + 1] - in.qoffs[J.query]);
        const int64_t n = (int64_t)J.win_hi - J.win_lo;
        BandJobDev d{};
        d.q_off = in.qoffs[J.query];
        d.s_off = in.soffs[J.subject] + J.win_lo;
        d.m = (uint32_t)m;
        d.n = (uint32_t)n;
        host_jobs[k] = d;
        if (m == 0 || n == 0 || J.band_hi < J.band_lo) continue;
        const int64_t lo = std::max<int64_t>(J.band_lo, -(m - 1));
        const int64_t hi = std::min<int64_t>(J.band_hi, n - 1);
        const int64_t w = hi - lo + 1;
        if (w <= 0) continue;
        const int64_t i_first = std::max<int64_t>(0, -hi);
        const int64_t i_last = std::min<int64_t>(m - 1, n - 1 - lo);
        if (i_last < i_first) continue;
        d.lo = (int32_t)lo;
        d.w = (int32_t)w;
        d.i_first = (int32_t)i_first;
        d.rows = (int32_t)(i_last - i_first + 1);
        host_cells[k] = band_cells(d);
        if (in.swap_ok) {
            // transposed: rows over the window, band [-hi, -lo] (already
            // inside [-(n - 1), m - 1] since lo >= -(m - 1) and hi <= n - 1)
            const int64_t t_first = std::max<int64_t>(0, lo), t_last = std::min<int64_t>(n - 1, m - 1 + hi);
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
        host_jobs[k] = d;
        rows_of[k] = d.rows;
        // The packed kernel keeps row indices in 16 bits and window columns
        // (times 8) in 32 bits.
        const bool fits = d.rows <= (int32_t)kMaxBandRows && m < kMaxPackedLen && n < kMaxPackedLen;
        const std::vector<int>& choice = in.ends ? c->choice_ends : c->choice;
        const int shape = (in.packed_ok && fits && w < (int64_t)choice.size()) ? choice[w] : -1;
        if (shape < 0) {
            generic.push_back((uint32_t)k);
            continue;
        }
        thr_shape[k] = shape;
        est_work += 0.5 * (d.rows + c->shapes[shape].L + 3) * c->shapes[shape].pair_cost();
    }
    lap("jobs");
    // Many-row jobs: a fixed row count, or the share of the estimated time.
    const double est_ms = est_work / kGpuLaneInstrPerMs;
    for (uint64_t k = 0; k < njobs; ++k) {
        const int shape = thr_shape[k];
        if (shape < 0) continue;
        const Shape& s = c->shapes[shape];
        bool lat;
        if (c->latency_rows < 0)
            lat = (rows_of[k] + s.L + 3) * s.row_cost() / kLaneInstrPerMs > kLatencyShare * est_ms;
        else
            lat = c->latency_rows > 0 && rows_of[k] > c->latency_rows;
        const int w = host_jobs[k].w;
        if (lat)
            bucket[2 * (in.ends ? c->choice_lat_ends : c->choice_lat)[w] + 1].push_back((uint32_t)k);
        else
            bucket[2 * shape].push_back((uint32_t)k);
    }
    // Pairs. A pair runs for its longer job's rows, so jobs are paired with
    // the next longest of their bucket. Jobs of one query that start on the
    // same query row are paired among themselves first: such a pair reads
    // the same query code on every row and uses the pair-table flavour. The
    // rest (different queries or first rows, odd ones out) pair by rows.
    // Work chunks of 64 lanes (64 / L pairs) of one shape and flavour are
    // claimed by the persistent kernel in decreasing order of their estimated
    // time (the longest pair's row steps x the shape's per-row cost), longest
    // processing time first.
    std::vector<uint32_t> pairs;
    std::vector<double> chunk_est;
    auto add_chunks = [&](const std::vector<uint32_t>& v, int shape, uint32_t flags) {
        if (v.empty()) return;
        const Shape& s = c->shapes[shape];
        const uint32_t per_chunk = 64u / (uint32_t)s.L;
        const uint64_t first_pair = pairs.size() / 2;
        for (size_t i = 0; i < v.size(); i += 2) {
            pairs.push_back(v[i]);
            pairs.push_back(i + 1 < v.size() ? v[i + 1] : kNone);
        }
        const uint64_t np = pairs.size() / 2 - first_pair;
        for (uint64_t p0 = 0; p0 < np; p0 += per_chunk) {
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(per_chunk, np - p0);
            const int steps = (rows_of[v[2 * p0]] + s.L - 1 + 3) & ~3;
            host_blocks.push_back(BandBlock{(uint32_t)shape, (uint32_t)(first_pair + p0), cnt, flags});
            chunk_est.push_back((double)steps * s.row_cost());
        }
    };
    lap("buckets");
    // sort keys: (query start, first row, rows descending) when they fit 64
    // bits (they do unless the query blob is beyond 2^33 residues), and the
    // rows descending alone
    uint64_t max_qoff = 0, max_ifirst = 0;
    for (uint64_t k = 0; k < njobs; ++k)
        if (thr_shape[k] >= 0) {
            max_qoff = std::max<uint64_t>(max_qoff, host_jobs[k].q_off);
            max_ifirst = std::max<uint64_t>(max_ifirst, (uint64_t)host_jobs[k].i_first);
        }
    const int bq = bits_of(max_qoff), bi = bits_of(max_ifirst), br = 16;
    const bool radix_ok = bq + bi + br <= 64;
    std::vector<uint64_t> key_group, key_rows(njobs, 0);
    if (radix_ok) key_group.assign(njobs, 0);
    for (uint64_t k = 0; k < njobs; ++k) {
        if (thr_shape[k] < 0) continue;
        const uint64_t rdesc = 65535u - (uint64_t)rows_of[k];
        key_rows[k] = rdesc;
        if (radix_ok)
            key_group[k] = (host_jobs[k].q_off << (bi + br)) | ((uint64_t)host_jobs[k].i_first << br) | rdesc;
    }
    for (size_t bk = 0; bk < bucket.size(); ++bk) {
        auto& v = bucket[bk];
        if (v.empty()) continue;
        const int shape = (int)(bk / 2);
        // group by (query start, first row), longest first within a group
        if (radix_ok) {
            radix_sort_by(v, key_group, bq + bi + br);
        } else {
            std::stable_sort(v.begin(), v.end(), [&](uint32_t x, uint32_t y) {
                const BandJobDev &ja = host_jobs[x], &jb = host_jobs[y];
                if (ja.q_off != jb.q_off) return ja.q_off < jb.q_off;
                if (ja.i_first != jb.i_first) return ja.i_first < jb.i_first;
                return rows_of[x] > rows_of[y];
            });
        }
        std::vector<uint32_t> shared, rest;
        shared.reserve(v.size());
        auto same_group = [&](uint32_t x, uint32_t y) {
            if (radix_ok) return (key_group[x] >> br) == (key_group[y] >> br);
            return host_jobs[x].q_off == host_jobs[y].q_off && host_jobs[x].i_first == host_jobs[y].i_first;
        };
        for (size_t i = 0; i < v.size();) {
            if (host_jobs[v[i]].swapped) {  // rows from the residue blob: never shared
                rest.push_back(v[i++]);
                continue;
            }
            size_t j = i;
            while (j < v.size() && !host_jobs[v[j]].swapped && same_group(v[j], v[i])) ++j;
            const size_t even = i + ((j - i) & ~(size_t)1);
            shared.insert(shared.end(), v.begin() + i, v.begin() + even);
            if (even < j) rest.push_back(v[even]);
            i = j;
        }
        // shared pairs in order of their longer job's rows (pairs stay intact)
        {
            std::vector<uint32_t> order(shared.size() / 2);
            std::vector<uint64_t> pk(order.size());
            for (size_t k = 0; k < order.size(); ++k) {
                order[k] = (uint32_t)k;
                pk[k] = key_rows[shared[2 * k]];
            }
            radix_sort_by(order, pk, br);
            std::vector<uint32_t> sorted(shared.size());
            for (size_t k = 0; k < order.size(); ++k) {
                sorted[2 * k] = shared[2 * order[k]];
                sorted[2 * k + 1] = shared[2 * order[k] + 1];
            }
            shared.swap(sorted);
        }
        radix_sort_by(rest, key_rows, br);
        add_chunks(shared, shape, kChunkShared);
        add_chunks(rest, shape, 0);
    }
    lap("pairs");
    if (pairs.size() / 2 >= 0xFFFFFFFFull) return fail(SABER_HIP_EINVAL, "too many job pairs for one batch");
    {
        // longest estimated chunk first; the estimates are integers (row
        // steps x the shape's integer per-row cost), ordered descending
        std::vector<uint32_t> order(host_blocks.size());
        for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
        double emax = 0.0;
        for (double e : chunk_est) emax = std::max(emax, e);
        std::vector<uint64_t> ek(order.size());
        for (size_t i = 0; i < order.size(); ++i) ek[i] = (uint64_t)(emax - chunk_est[i]);
        radix_sort_by(order, ek, bits_of((uint64_t)emax));
        std::vector<BandBlock> sorted(order.size());
        for (size_t i = 0; i < order.size(); ++i) sorted[i] = host_blocks[order[i]];
        host_blocks.swap(sorted);
    }

    lap("chunk order");

    if (int r = upload(b->jobs, host_jobs.data(), njobs, st)) return r;
    if (int r = upload(b->cells, host_cells.data(), njobs, st)) return r;
    if (int r = upload(b->pairs, pairs.data(), pairs.size(), st)) return r;
    if (int r = upload(b->blocks, host_blocks.data(), host_blocks.size(), st)) return r;
    out->npairs = pairs.size() / 2;
    out->nchunks = host_blocks.size();
    out->ngeneric = (uint32_t)generic.size();
    out->npacked = (uint64_t)std::count_if(thr_shape.begin(), thr_shape.end(), [](int s) { return s >= 0; });
    // int32 jobs: ids and work offsets
    if (!generic.empty()) {
        std::vector<uint64_t> off(generic.size());
        uint64_t tot = 0;
        for (size_t i = 0; i < generic.size(); ++i) {
            off[i] = tot;
            tot += 4 * ((uint64_t)host_jobs[generic[i]].w + 1);
        }
        if (int r = upload(b->gen_ids, generic.data(), generic.size(), st)) return r;
        if (int r = upload(b->gen_off, off.data(), off.size(), st)) return r;
        out->gen_work = tot;
    }
    out->wmax = 0;
    for (uint64_t k = 0; k < njobs; ++k)
        if (thr_shape[k] >= 0) out->wmax = std::max(out->wmax, host_jobs[k].w);
    // the host arrays stay alive until their copies are done
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

// A band batch over host inputs (queries / subjects / jobs, the public entry
// point) or over inputs already on the device (dev.jobs set: jobs, query
// offsets, codes and biases in device memory, subjects from the context
// database, planned on the device). The caller holds c->mu.
int saber::band_batch_create_impl(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                                  const saber_hip_band_job* jobs, uint64_t njobs, const DevBandInput* dev,
                                  const saber_hip_scheme* scheme, int ends, saber_hip_band_batch** out) {
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t* soffs = subjects ? subjects->offsets : c->db_off.data();
    const uint32_t nsub = subjects ? subjects->count : (c->db_off.empty() ? 0 : (uint32_t)c->db_off.size() - 1);

    auto b = std::make_unique<saber_hip_band_batch>();
    b->ctx = c;
    b->njobs = njobs;
    b->has_bias = dev ? dev->qbias != nullptr : queries->bias != nullptr;
    b->ends = ends != 0;
    b->go = scheme->gap_open;
    b->ge = scheme->gap_extend;

    // Can the packed int16 kernel be exact for this scheme? (Gap costs within
    // the range GapCosts::from_user accepts, substitution scores within int8.)
    bool packed_ok = !(c->flags & SABER_HIP_FORCE_EXACT32);
    packed_ok = packed_ok && scheme->gap_open >= 0 && scheme->gap_open <= 1000 &&
                scheme->gap_extend >= 0 && scheme->gap_extend <= 1000;
    for (int i = 0; i < 576; ++i)
        if (scheme->matrix[i] < -128 || scheme->matrix[i] > 127) packed_ok = false;

    // Scores-only jobs without bias may run transposed (query and window
    // exchanged, band negated: the same score under a symmetric matrix) when
    // that gives them fewer rows, as it does for windows narrower than their
    // band, whose band otherwise sweeps mostly outside the rectangle.
    bool swap_ok = packed_ok && !b->ends && !b->has_bias;
    for (int x = 0; x < 24 && swap_ok; ++x)
        for (int y = 0; y < x; ++y)
            if (scheme->matrix[x * 24 + y] != scheme->matrix[y * 24 + x]) swap_ok = false;

    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "#   plan %s %.1f ms\n", what, std::chrono::duration<double, std::milli>(t1 - tp).count());
        tp = t1;
    };

    // Plan: on the device for large batches, else on the host (the same plan).
    hipStream_t st = c->stream;
    BandPlanInput in{jobs, njobs, dev ? nullptr : queries->offsets, dev ? dev->qcount : queries->count, soffs,
                     subjects ? nullptr : c->db_off_dev.p, nsub, packed_ok, swap_ok};
    in.ends = b->ends;
    if (dev) {
        in.jobs_dev = dev->jobs;
        in.qoffs_dev = dev->qoffs;
    }
    const bool on_device = dev || (c->flags & SABER_HIP_DEVICE_PLAN) ||
                           (!(c->flags & SABER_HIP_HOST_PLAN) && njobs >= kDevicePlanMinJobs);
    BandPlan plan;
    if (on_device) {
        if (int r = band_plan_device(c, st, in, b->jobs, b->cells, b->pairs, b->blocks, b->gen_ids, b->gen_off, &plan))
            return r;
        lap("on the device");
    } else {
        if (int r = plan_on_host(c, st, in, b.get(), &plan, lap)) return r;
    }
    b->nchunks = plan.nchunks;
    b->ngeneric = plan.ngeneric;
    b->npacked = plan.npacked;

    // Score table for the packed kernel: int8 T[a][c] = S(a, c) for codes
    // a, c < 24 (stride 25), column 24 = sentinel and row 24 = dead row
    // (-128). Only built when every score fits int8 (packed_ok); otherwise no
    // job uses it.
    std::vector<uint32_t> table(kTable8Words, 0);
    {
        int8_t* t8 = reinterpret_cast<int8_t*>(table.data());
        for (int a = 0; a <= kBandSentinel; ++a)
            for (int cc = 0; cc <= kBandSentinel; ++cc) {
                const int v = (a == kBandSentinel || cc == kBandSentinel)
                                  ? -128
                                  : std::max(-128, std::min(127, scheme->matrix[a * 24 + cc]));
                t8[a * kTable8Stride + cc] = (int8_t)v;
            }
    }
    const std::vector<uint32_t> ptable = build_pair_table(scheme->matrix);

    uint64_t nq = 0;
    if (dev) {
        nq = dev->qlen_total;
        b->qres = dev->qres;
        b->qbias = dev->qbias;
    } else {
        nq = queries->count ? queries->offsets[queries->count] : 0;
        if (int r = upload_codes(b->q, queries->residues, nq, st)) return r;
        b->qres = b->q.p + kGuard;
        if (b->has_bias) {
            if (int r = upload_guarded(b->bias, queries->bias, nq, st)) return r;
            b->qbias = b->bias.p + kGuard;
        }
    }
    if (subjects) {
        const uint64_t ns = subjects->count ? subjects->offsets[subjects->count] : 0;
        if (int r = upload_codes(b->s, subjects->residues, ns, st)) return r;
        b->sres = b->s.p + kGuard;
        b->slen_total = ns;
    } else {
        if (!c->db.p && njobs > 0) return fail(SABER_HIP_ESTATE, "no database uploaded");
        b->sres = c->db.p + kGuard;
        b->slen_total = c->db_off.empty() ? 0 : c->db_off.back();
    }
    b->qlen_total = nq;
    if (int r = upload(b->table, table.data(), table.size(), st)) return r;
    if (int r = upload(b->pair_table, ptable.data(), ptable.size(), st)) return r;
    HIP_TRY(b->next_chunk.alloc(1));
    if (int r = upload(b->mat, scheme->matrix, 576, st)) return r;
    HIP_TRY(b->flagged.alloc(njobs ? njobs : 1));
    HIP_TRY(b->nflagged.alloc(1));
    if (plan.ngeneric) HIP_TRY(b->gen_work.alloc(plan.gen_work));
    // saturated-job slots: any packed job may saturate, so a slot holds the
    // widest packed job's two band rows
    if (njobs > 0) {
        b->flag_slot = 4u * ((uint32_t)plan.wmax + 1u);
        b->flag_threads = (uint32_t)std::min<uint64_t>(8192, std::max<uint64_t>(128, (njobs / 64 + 127) / 128 * 128));
        HIP_TRY(b->flag_work.alloc((uint64_t)b->flag_threads * b->flag_slot));
    }
    HIP_TRY(b->score.alloc(njobs ? njobs : 1));
    HIP_TRY(b->q_end.alloc(njobs ? njobs : 1));
    HIP_TRY(b->s_end.alloc(njobs ? njobs : 1));
    HIP_TRY(hipStreamSynchronize(st));
    lap("uploads");
    *out = b.release();
    return SABER_HIP_OK;
}

extern "C" {

int saber_hip_band_batch_create(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                                uint64_t njobs, const saber_hip_scheme* scheme, int ends,
                                saber_hip_band_batch** out) {
    if (int r_ = injected(c)) return r_;
    if (!c || !out || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (njobs > 0 && !jobs) return fail(SABER_HIP_EINVAL, "jobs is null");
    if (njobs >= (uint64_t)kNone) return fail(SABER_HIP_EINVAL, "too many jobs for one batch");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    std::lock_guard<std::mutex> g(c->mu);
    return band_batch_create_impl(c, queries, subjects, jobs, njobs, nullptr, scheme, ends, out);
}

static GenericArgs generic_args(saber_hip_band_batch* b) {
    GenericArgs a{};
    a.jobs = b->jobs.p;
    a.qres = b->qres;
    a.qbias = b->has_bias ? b->qbias : nullptr;
    a.sres = b->sres;
    a.mat = b->mat.p;
    a.go = b->go;
    a.ge = b->ge;
    a.out = BandOutputs{b->score.p, b->q_end.p, b->s_end.p, b->cells.p};
    return a;
}

// One scoring pass, entirely on the device (so it can be captured as a
// graph): output clear, the packed kernel, the int32 kernel over the jobs
// routed to it at plan time and over the packed kernel's saturated jobs (their
// count is read on the device). ev_begin / ev_end bracket the packed kernel.
static int enqueue_step(saber_hip_band_batch* b, hipEvent_t ev_begin, hipEvent_t ev_end) {
    hipStream_t st = b->ctx->stream;
    const uint64_t n = b->njobs;
    // Every packed and int32 job's outputs are written by its kernel (a
    // saturated packed job's by the rescore); only jobs without band cells
    // need the zeros.
    if (b->npacked + b->ngeneric < n) {
        HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(b->score.p), n, 0, st));
        HIP_TRY(fill_u32(b->q_end.p, n, 0, st));
        HIP_TRY(fill_u32(b->s_end.p, n, 0, st));
    }
    HIP_TRY(fill_u32(b->nflagged.p, 1, 0, st));
    HIP_TRY(fill_u32(b->next_chunk.p, 1, 0, st));
    BandKernelArgs a{};
    a.pairs = b->pairs.p;
    a.npairs = (uint32_t)(b->pairs.n / 2);
    a.blocks = b->blocks.p;
    a.nchunks = (uint32_t)b->nchunks;
    a.next_chunk = b->next_chunk.p;
    a.pair_table = b->pair_table.p;
    a.jobs = b->jobs.p;
    a.qres = b->qres;
    a.qbias = b->has_bias ? b->qbias : nullptr;
    a.sres = b->sres;
    a.qres_len = b->qlen_total;
    a.sres_len = b->slen_total;
    a.table = b->table.p;
    a.go = b->go;
    a.ge = b->ge;
    a.out = BandOutputs{b->score.p, b->q_end.p, b->s_end.p, b->cells.p};
    a.flagged = b->flagged.p;
    a.nflagged = b->nflagged.p;
    if (ev_begin) HIP_TRY(hipEventRecord(ev_begin, st));
    // one persistent workgroup per CU (fewer when there is little work)
    const uint32_t wg = (uint32_t)std::min<uint64_t>((uint64_t)b->ctx->cus, (b->nchunks + 11) / 12);
    if (a.nchunks) HIP_TRY(launch_band_multi(a, wg, b->has_bias, b->ends, st));
    if (ev_end) HIP_TRY(hipEventRecord(ev_end, st));
    if (b->ngeneric) {
        GenericArgs g = generic_args(b);
        g.ids = b->gen_ids.p;
        g.nids = b->ngeneric;
        g.work = b->gen_work.p;
        g.work_off = b->gen_off.p;
        HIP_TRY(launch_generic(g, st));
    }
    if (b->flag_threads) {
        GenericArgs g = generic_args(b);
        g.ids = b->flagged.p;
        g.work = b->flag_work.p;
        HIP_TRY(launch_generic_slots(g, b->nflagged.p, b->flag_threads, b->flag_slot, st));
    }
    return SABER_HIP_OK;
}

// Enqueue one pass: the first pass directly (a batch run once, as by
// saber_hip_score_bands, never pays for a capture), later passes from the
// graph captured on the second, unless graphs are switched off for the
// context.
static int enqueue_run(saber_hip_band_batch* b) {
    hipStream_t st = b->ctx->stream;
    if ((b->ctx->flags & SABER_HIP_NO_GRAPH) || b->passes++ == 0) return enqueue_step(b, nullptr, nullptr);
    if (!b->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int r = enqueue_step(b, nullptr, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (r != SABER_HIP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, st));
    return SABER_HIP_OK;
}

}  // extern "C"

int saber::band_batch_enqueue(saber_hip_band_batch* b) {
    if (int r = enqueue_run(b)) return r;
    b->ran = true;
    return SABER_HIP_OK;
}
const int32_t* saber::band_batch_scores(const saber_hip_band_batch* b) { return b->score.p; }
const uint64_t* saber::band_batch_cells(const saber_hip_band_batch* b) { return b->cells.p; }

extern "C" {

int saber_hip_band_batch_run(saber_hip_band_batch* b) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (int r = enqueue_run(b)) return r;
    b->ran = true;
    return SABER_HIP_OK;
}

int saber_hip_band_batch_wait(saber_hip_band_batch* b) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return SABER_HIP_OK;
}

int saber_hip_band_batch_read(saber_hip_band_batch* b, int32_t* score, uint32_t* q_end,
                              uint32_t* s_end, uint64_t* cells) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    if (!b->ran) return fail(SABER_HIP_ESTATE, "batch has not run");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const uint64_t n = b->njobs;
    if (!b->ends && (q_end || s_end)) return fail(SABER_HIP_ESTATE, "batch computes scores only");
    if (n == 0) return SABER_HIP_OK;
    if (score) HIP_TRY(hipMemcpyAsync(score, b->score.p, 4 * n, hipMemcpyDeviceToHost, st));
    if (q_end) HIP_TRY(hipMemcpyAsync(q_end, b->q_end.p, 4 * n, hipMemcpyDeviceToHost, st));
    if (s_end) HIP_TRY(hipMemcpyAsync(s_end, b->s_end.p, 4 * n, hipMemcpyDeviceToHost, st));
    if (cells) HIP_TRY(hipMemcpyAsync(cells, b->cells.p, 8 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

void saber_hip_band_batch_destroy(saber_hip_band_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    delete b;
}

int saber_hip_band_batch_time(saber_hip_band_batch* b, int warmup, int iters, double* ms_run,
                              double* band_ms, uint64_t* variant_cells, int* nvar,
                              uint64_t* cells) {
    if (!b || iters < 1 || warmup < 0) return fail(SABER_HIP_EINVAL, "bad arguments");
    int nv = 0;
    band_variants(&nv);
    const int cap = nvar ? std::min(*nvar, nv) : 0;
    for (int i = 0; i < warmup; ++i)
        if (int r = saber_hip_band_batch_run(b)) return r;
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    std::vector<hipEvent_t> eb(iters), ee(iters);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (auto& e : eb) HIP_TRY(hipEventCreate(&e));
    for (auto& e : ee) HIP_TRY(hipEventCreate(&e));
    // whole passes as the caller runs them (the captured graph)
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it)
        if (int r = enqueue_run(b)) return r;
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    b->ran = true;
    float tot = 0.f;
    HIP_TRY(hipEventElapsedTime(&tot, e0, e1));
    if (ms_run) *ms_run = tot / iters;
    HIP_TRY(hipMemcpy(&b->last_flagged, b->nflagged.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (band_ms) {
        // the packed kernel alone: the same passes enqueued directly, with
        // events around its launch
        for (int it = 0; it < iters; ++it)
            if (int r = enqueue_step(b, eb[it], ee[it])) return r;
        HIP_TRY(hipStreamSynchronize(st));
        double sum = 0.0;
        for (int it = 0; it < iters; ++it) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, eb[it], ee[it]));
            sum += ms;
        }
        *band_ms = sum / iters;
    }
    if (nvar) *nvar = nv;
    for (auto& e : eb) (void)hipEventDestroy(e);
    for (auto& e : ee) (void)hipEventDestroy(e);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (cells || variant_cells) {
        std::vector<uint64_t> c(b->njobs);
        if (b->njobs) HIP_TRY(hipMemcpy(c.data(), b->cells.p, 8 * b->njobs, hipMemcpyDeviceToHost));
        uint64_t t = 0;
        for (uint64_t x : c) t += x;
        if (cells) *cells = t;
        if (variant_cells) {
            std::vector<uint32_t> pairs(b->pairs.n);
            if (!pairs.empty()) HIP_TRY(hipMemcpy(pairs.data(), b->pairs.p, 4 * pairs.size(), hipMemcpyDeviceToHost));
            std::vector<BandBlock> blocks(b->nchunks);
            if (!blocks.empty())
                HIP_TRY(hipMemcpy(blocks.data(), b->blocks.p, sizeof(BandBlock) * blocks.size(), hipMemcpyDeviceToHost));
            for (int v = 0; v < cap; ++v) variant_cells[v] = 0;
            for (const BandBlock& k : blocks) {
                const int v = 2 * (int)k.shape + (b->has_bias ? 1 : 0);
                if (v >= cap) continue;
                for (uint64_t x = 2 * (uint64_t)k.pair_off; x < 2 * ((uint64_t)k.pair_off + k.npairs); ++x)
                    if (pairs[x] != kNone) variant_cells[v] += c[pairs[x]];
            }
        }
    }
    return SABER_HIP_OK;
}

int saber_hip_band_batch_plan(saber_hip_band_batch* b, uint64_t* npairs, uint64_t* nchunks, uint32_t* pairs,
                              uint32_t* blocks, int32_t* geometry) {
    if (!b || !npairs || !nchunks) return fail(SABER_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    *npairs = b->nchunks ? b->pairs.n / 2 : 0;
    *nchunks = b->nchunks;
    if (pairs && *npairs) HIP_TRY(hipMemcpy(pairs, b->pairs.p, 8 * *npairs, hipMemcpyDeviceToHost));
    if (blocks && b->nchunks) HIP_TRY(hipMemcpy(blocks, b->blocks.p, sizeof(BandBlock) * b->nchunks, hipMemcpyDeviceToHost));
    if (geometry && b->njobs) {
        std::vector<BandJobDev> j(b->njobs);
        HIP_TRY(hipMemcpy(j.data(), b->jobs.p, sizeof(BandJobDev) * b->njobs, hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < b->njobs; ++k) {
            int32_t* o = geometry + 5 * k;
            o[0] = j[k].lo;
            o[1] = j[k].w;
            o[2] = j[k].i_first;
            o[3] = j[k].rows;
            o[4] = (int32_t)j[k].swapped;
        }
    }
    return SABER_HIP_OK;
}

int saber_hip_score_bands(saber_hip_ctx* c, const saber_hip_seqset* queries,
                          const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                          uint64_t njobs, const saber_hip_scheme* scheme, int32_t* score,
                          uint32_t* q_end, uint32_t* s_end, uint64_t* cells) {
    if (int r_ = injected(c)) return r_;
    if (njobs > 0 && !score) return fail(SABER_HIP_EINVAL, "score is null");
    saber_hip_band_batch* b = nullptr;
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "# score_bands %llu jobs: %s %.1f ms\n", (unsigned long long)njobs, what,
                     std::chrono::duration<double, std::milli>(t1 - t0).count());
        t0 = t1;
    };
    if (int r = saber_hip_band_batch_create(c, queries, subjects, jobs, njobs, scheme, q_end || s_end, &b)) return r;
    lap("plan + upload");
    int r = saber_hip_band_batch_run(b);
    if (r == SABER_HIP_OK) r = saber_hip_band_batch_wait(b);
    lap("run");
    if (r == SABER_HIP_OK) r = saber_hip_band_batch_read(b, score, q_end, s_end, cells);
    lap("read");
    saber_hip_band_batch_destroy(b);
    lap("destroy");
    return r;
}

}  // extern "C"

namespace {

// Host staging that grows to the largest request and is not initialised
// (every byte read back was first written by a copy or the caller).
template <class T>
struct HostBuf {
    std::unique_ptr<T[]> p;
    uint64_t n = 0;
    T* get(uint64_t count) {
        if (count > n) {
            p.reset(new T[count]);
            n = count;
        }
        return p.get();
    }
};

// The residues (and bias) of one band_align call on the device and its
// scratch (work rows, direction bytes, operations), kept by a caller that
// makes several calls over the same sequence sets: the residues are uploaded
// once, the scratch grows to the largest call.
struct AlignSeqs {
    DevBuf<uint8_t> q, sbuf;
    DevBuf<int8_t> bias;
    bool ready = false;
    DevBuf<int32_t> work;
    DevBuf<uint8_t> dir, codes, dcodes;
    DevBuf<uint32_t> lens, dlens;
    DevBuf<uint64_t> dst_off;
    HostBuf<uint8_t> hcodes;
    HostBuf<uint32_t> hlens;
    // per-call job data, refilled by every call of a caller's series
    DevBuf<BandJobDev> djobs;
    DevBuf<int32_t> mat, stops;
    DevBuf<AlignRecord> recs;
    DevBuf<uint32_t> dids, dorder;
    DevBuf<uint64_t> dwoff, ddoff, dooff;
};

// saber_hip_band_align; iters > 0 also times iters repetitions of every
// kernel launch with HIP events (the kernel is idempotent) and adds their mean
// to *kernel_ms.

int band_align_impl(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                    const saber_hip_band_job* jobs, const int32_t* stop_at, uint64_t njobs,
                    const saber_hip_scheme* scheme, uint64_t max_dir_bytes, saber_hip_alignment* out,
                    uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity, int iters, double* kernel_ms,
                    AlignSeqs* seqs = nullptr) {
    if (kernel_ms) *kernel_ms = 0.0;
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (njobs == 0) return SABER_HIP_OK;
    if (!jobs || !out) return fail(SABER_HIP_EINVAL, "null argument");
    if (njobs >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "too many jobs");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint64_t* soffs = subjects ? subjects->offsets : c->db_off.data();
    const uint32_t nsub = subjects ? subjects->count : (c->db_off.empty() ? 0 : (uint32_t)c->db_off.size() - 1);
    if (!subjects && !c->db.p) return fail(SABER_HIP_ESTATE, "no database uploaded");
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        (void)hipStreamSynchronize(st);
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "#   align %s %.2f ms\n", what, std::chrono::duration<double, std::milli>(t1 - tp).count());
        tp = t1;
    };
    // Jobs: clipped bands (Band::clip, banded.rs:45-47), the reference's
    // memory check on m * w direction bytes, output operation offsets.
    std::vector<BandJobDev> dj(njobs);
    std::vector<uint64_t> op_off(njobs);
    std::vector<uint8_t> capped(njobs, 0);
    uint64_t ops_total = 0;
    for (uint64_t k = 0; k < njobs; ++k) {
        const saber_hip_band_job& J = jobs[k];
        if (J.query >= queries->count) return fail(SABER_HIP_EINVAL, "job query index out of range");
        if (J.subject >= nsub) return fail(SABER_HIP_EINVAL, "job subject index out of range");
        const uint64_t slen = soffs[J.subject + 1] - soffs[J.subject];
        if (J.win_hi < J.win_lo || J.win_hi > slen) return fail(SABER_HIP_EINVAL, "job window outside its subject");
        const int64_t m = (int64_t)(queries->offsets[J.query + 1] - queries->offsets[J.query]);
        const int64_t n = (int64_t)J.win_hi - J.win_lo;
        BandJobDev d{};
        d.q_off = queries->offsets[J.query];
        d.s_off = soffs[J.subject] + J.win_lo;
        d.m = (uint32_t)m;
        d.n = (uint32_t)n;
        op_off[k] = ops_total;
        ops_total += (uint64_t)(m + n);
        if (m > 0 && n > 0 && J.band_hi >= J.band_lo) {
            const int64_t lo = std::max<int64_t>(J.band_lo, -(m - 1));
            const int64_t hi = std::min<int64_t>(J.band_hi, n - 1);
            const int64_t w = hi - lo + 1;
            if (w > 0) {
                if ((uint64_t)m * (uint64_t)w > max_dir_bytes) {
                    capped[k] = 1;
                } else {
                    const int64_t i_first = std::max<int64_t>(0, -hi);
                    const int64_t i_last = std::min<int64_t>(m - 1, n - 1 - lo);
                    d.lo = (int32_t)lo;
                    d.w = (int32_t)w;
                    d.i_first = (int32_t)i_first;
                    d.rows = (int32_t)(i_last >= i_first ? i_last - i_first + 1 : 0);
                }
            }
        }
        dj[k] = d;
    }
    if (ops_total > op_capacity || (ops_total > 0 && (!op_codes || !op_lens)))
        return fail(SABER_HIP_EINVAL, "operation buffers need the sum over jobs of query + window length entries");
    AlignSeqs local;
    AlignSeqs& sq = seqs ? *seqs : local;
    DevBuf<uint8_t>& q = sq.q;
    DevBuf<int8_t>& bias = sq.bias;
    DevBuf<BandJobDev>& djobs = sq.djobs;
    DevBuf<int32_t>& mat = sq.mat;
    DevBuf<int32_t>& stops = sq.stops;
    DevBuf<AlignRecord>& recs = sq.recs;
    if (!sq.ready) {
        const uint64_t nqres = queries->offsets[queries->count];
        if (int r = upload_codes(q, queries->residues, nqres, st)) return r;
        if (queries->bias)
            if (int r = upload_guarded(bias, queries->bias, nqres, st)) return r;
        if (subjects)
            if (int r = upload_codes(sq.sbuf, subjects->residues, subjects->offsets[subjects->count], st)) return r;
        sq.ready = true;
    }
    const uint8_t* sres = subjects ? sq.sbuf.p + kGuard : c->db.p + kGuard;
    lap("jobs + sequences");
    if (int r = upload_keep(djobs, dj.data(), njobs, st)) return r;
    if (int r = upload_keep(mat, scheme->matrix, 576, st)) return r;
    if (stop_at)
        if (int r = upload_keep(stops, stop_at, njobs, st)) return r;
    HIP_TRY(recs.reserve(njobs));
    // Launches of jobs whose scratch (work rows, direction bytes, operation
    // bytes) fits a budget: the lane classes (band widths up to 64 x 17) in
    // one kind of launch, wider jobs on one thread each with their rows in
    // the work buffer in the other.
    uint64_t kBudget = 2048ull << 20;
    if (const char* e = std::getenv("SABER_HIP_ALIGN_SCRATCH_MB")) {
        const long long v = std::strtoll(e, nullptr, 10);
        if (v > 0 && v < (1ll << 24)) kBudget = (uint64_t)v << 20;
    }
    std::vector<AlignRecord> hrec(njobs);
    std::vector<uint32_t> ids;
    std::vector<uint64_t> woff, doff, ooff;
    DevBuf<uint32_t>& dids = sq.dids;
    DevBuf<uint32_t>& dorder = sq.dorder;
    std::vector<uint32_t> border;
    std::vector<uint64_t> bkey;
    DevBuf<uint64_t>& dwoff = sq.dwoff;
    DevBuf<uint64_t>& ddoff = sq.ddoff;
    DevBuf<uint64_t>& dooff = sq.dooff;
    DevBuf<int32_t>& work = sq.work;
    DevBuf<uint8_t>& dir = sq.dir;
    DevBuf<uint8_t>& codes = sq.codes;
    DevBuf<uint32_t>& lens = sq.lens;
    HostBuf<uint8_t>& hcodes = sq.hcodes;
    HostBuf<uint32_t>& hlens = sq.hlens;
    // Lane class of a job: the least lane count of 16 cells that covers the
    // band width (kAlignLaneClasses: one thread per job). A job whose row
    // steps on those lanes (rows + L - 1) exceed both kAlignMinLongSteps and
    // the batch's mean load of a SIMD lane -- the sum over jobs of steps x
    // lanes over the device's 1024 SIMDs x 64 lanes -- would finish after the
    // rest: it takes the least 4-cell class that covers its width, if any.
    // Runnable jobs in class order, longest first within a class.
    auto class_in = [&](int32_t w, int first, int last) {
        for (int x = first; x < last; ++x)
            if (w <= align_class_lanes(x) * align_class_cells(x)) return x;
        return kAlignLaneClasses;
    };
    std::vector<uint32_t> order;
    std::vector<uint8_t> cls(njobs, 0);
    uint64_t lane_steps = 0;
    for (uint64_t j = 0; j < njobs; ++j) {
        const BandJobDev& d = dj[j];
        if (capped[j] || d.w <= 0 || d.rows <= 0) continue;
        cls[j] = (uint8_t)class_in(d.w, 0, kAlignWideClasses);
        if (cls[j] < kAlignLaneClasses) {
            const uint64_t L = (uint64_t)align_class_lanes(cls[j]);
            lane_steps += ((uint64_t)d.rows + L - 1) * L;
        }
        order.push_back((uint32_t)j);
    }
    const uint64_t long_steps = std::max<uint64_t>(kAlignMinLongSteps, lane_steps / (1024 * 64));
    for (const uint32_t j : order) {
        if (cls[j] >= kAlignLaneClasses) continue;
        const BandJobDev& d = dj[j];
        if ((uint64_t)d.rows + (uint64_t)align_class_lanes(cls[j]) - 1 > long_steps) {
            const int x = class_in(d.w, kAlignWideClasses, kAlignLaneClasses);
            if (x < kAlignLaneClasses) cls[j] = (uint8_t)x;
        }
    }
    // the lane kernel's 32-bit end-cell key must cover every row
    for (const uint32_t j : order)
        if (cls[j] < kAlignLaneClasses && !align_lane_key_fits((uint64_t)dj[j].rows, cls[j]))
            cls[j] = (uint8_t)kAlignLaneClasses;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return cls[x] != cls[y] ? cls[x] < cls[y] : dj[x].rows > dj[y].rows;
    });
    size_t pos = 0;
    uint64_t ops_out = 0;   // operations written to the caller's buffers
    {
        while (pos < order.size()) {
            ids.clear();
            woff.clear();
            doff.clear();
            ooff.clear();
            uint64_t wt = 0, dt = 0, ot = 0;
            const bool lanes = cls[order[pos]] < kAlignLaneClasses;
            for (; pos < order.size(); ++pos) {
                const uint32_t k = order[pos];
                const BandJobDev& d = dj[k];
                if ((cls[k] < kAlignLaneClasses) != lanes) break;
                // direction rows of L C bytes on L lanes of C cells (one
                // store per lane and row, one spare row for the lag steps), w
                // bytes on one thread
                const uint64_t wb = lanes ? 0 : 8ull * ((uint64_t)d.w + 1),
                               db = lanes ? ((uint64_t)d.rows + 1) *
                                                (uint64_t)(align_class_lanes(cls[k]) * align_class_cells(cls[k]))
                                          : (uint64_t)d.rows * (uint64_t)d.w,
                               ob = (uint64_t)d.m + d.n;
                if (!ids.empty() && (wt + dt + 5 * ot + wb + db + 5 * ob > kBudget || ids.size() >= (1u << 20)))
                    break;
                ids.push_back(k);
                woff.push_back(wt / 4);
                doff.push_back(dt);
                ooff.push_back(ot);
                wt += wb;
                dt += db;
                ot += ob;
            }
            AlignLaneClasses lc{};
            if (lanes) {
                uint32_t slot = 0, blk = 0;
                for (int x = 0; x < kAlignLaneClasses; ++x) {
                    lc.slot_start[x] = slot;
                    lc.block_start[x] = blk;
                    uint32_t cnt = 0;
                    while (slot + cnt < ids.size() && cls[ids[slot + cnt]] == x) ++cnt;
                    slot += cnt;
                    blk += (uint32_t)(((uint64_t)cnt * (uint64_t)align_class_lanes(x) + 255) / 256);
                }
                lc.slot_start[kAlignLaneClasses] = slot;
                lc.block_start[kAlignLaneClasses] = blk;
                // Blocks by the time of their first (longest) job, rows +
                // L - 1 steps of C cells a lane, longest first across the
                // classes.
                border.clear();
                bkey.clear();
                for (int x = 0; x < kAlignLaneClasses; ++x) {
                    const uint32_t lanes_x = (uint32_t)align_class_lanes(x), per = 256u / lanes_x;
                    for (uint32_t b = 0; b < lc.block_start[x + 1] - lc.block_start[x]; ++b) {
                        const uint32_t first = ids[lc.slot_start[x] + b * per];
                        border.push_back((uint32_t)x << 24 | b);
                        bkey.push_back(((uint64_t)dj[first].rows + lanes_x - 1) * (uint64_t)align_class_cells(x));
                    }
                }
                std::vector<uint32_t> bidx(border.size());
                for (uint32_t b = 0; b < bidx.size(); ++b) bidx[b] = b;
                std::stable_sort(bidx.begin(), bidx.end(), [&](uint32_t u, uint32_t v) { return bkey[u] > bkey[v]; });
                std::vector<uint32_t> sorted(bidx.size());
                for (size_t b = 0; b < bidx.size(); ++b) sorted[b] = border[bidx[b]];
                border.swap(sorted);
                if (int r = upload_keep(dorder, border.data(), border.size(), st)) return r;
                lc.order = dorder.p;
            }
            if (int r = upload_keep(dids, ids.data(), ids.size(), st)) return r;
            if (int r = upload_keep(dwoff, woff.data(), woff.size(), st)) return r;
            if (int r = upload_keep(ddoff, doff.data(), doff.size(), st)) return r;
            if (int r = upload_keep(dooff, ooff.data(), ooff.size(), st)) return r;
            if (wt && work.n < wt / 4) HIP_TRY(work.alloc(wt / 4));
            if (dir.n < dt) HIP_TRY(dir.alloc(dt));
            if (codes.n < ot) {
                HIP_TRY(codes.alloc(ot));
                HIP_TRY(lens.alloc(ot));
            }
            AlignArgs a{};
            a.ids = dids.p;
            a.n = (uint32_t)ids.size();
            a.jobs = djobs.p;
            a.qres = q.p + kGuard;
            a.qbias = queries->bias ? bias.p + kGuard : nullptr;
            a.sres = sres;
            a.mat = mat.p;
            a.go = scheme->gap_open;
            a.ge = scheme->gap_extend;
            a.stop_at = stop_at ? stops.p : nullptr;
            a.work_off = dwoff.p;
            a.work = work.p;
            a.dir_off = ddoff.p;
            a.dir = dir.p;
            a.op_off = dooff.p;
            a.op_codes = codes.p;
            a.op_lens = lens.p;
            a.out = recs.p;
            lap("plan + uploads");
            auto launch = [&]() { return lanes ? launch_band_align_lanes(a, lc, st) : launch_band_align(a, st); };
            HIP_TRY(launch());
            lap("kernel");
            if (iters > 0 && kernel_ms) {
                hipEvent_t e0, e1;
                HIP_TRY(hipEventCreate(&e0));
                HIP_TRY(hipEventCreate(&e1));
                HIP_TRY(hipEventRecord(e0, st));
                for (int it = 0; it < iters; ++it) HIP_TRY(launch());
                HIP_TRY(hipEventRecord(e1, st));
                HIP_TRY(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                *kernel_ms += ms / iters;
            }
            // Records first; then only the operations the alignments use,
            // packed densely on the device (the slots hold m + n entries
            // each, mostly unused).
            HIP_TRY(hipMemcpyAsync(hrec.data(), recs.p, sizeof(AlignRecord) * njobs, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            std::vector<uint64_t> dense(ids.size());
            uint64_t used = 0;
            for (size_t x = 0; x < ids.size(); ++x) {
                const AlignRecord& r = hrec[ids[x]];
                dense[x] = used;
                used += r.status == 0 ? r.nops : 0;
            }
            uint8_t* hc = hcodes.get(std::max<uint64_t>(used, 1));
            uint32_t* hl = hlens.get(std::max<uint64_t>(used, 1));
            if (used) {
                if (int r = upload_keep(sq.dst_off, dense.data(), dense.size(), st)) return r;
                if (sq.dcodes.n < used) {
                    HIP_TRY(sq.dcodes.alloc(used));
                    HIP_TRY(sq.dlens.alloc(used));
                }
                HIP_TRY(launch_gather_ops((uint32_t)ids.size(), dids.p, recs.p, dooff.p, sq.dst_off.p, codes.p,
                                          lens.p, sq.dcodes.p, sq.dlens.p, st));
                HIP_TRY(hipMemcpyAsync(hc, sq.dcodes.p, used, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipMemcpyAsync(hl, sq.dlens.p, 4 * used, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
            }
            lap("copy back");
            // packed into the caller's buffers, launch after launch
            if (used) {
                std::memcpy(op_codes + ops_out, hc, used);
                std::memcpy(op_lens + ops_out, hl, 4ull * used);
            }
            for (size_t x = 0; x < ids.size(); ++x) {
                const uint32_t id = ids[x];
                const AlignRecord& r = hrec[id];
                saber_hip_alignment& o = out[id];
                o = saber_hip_alignment{};
                o.status = r.status;
                o.score = r.score;
                o.q_start = r.q_start;
                o.q_end = r.q_end;
                o.s_start = r.s_start;
                o.s_end = r.s_end;
                o.nops = r.status == 0 ? r.nops : 0;
                o.op_off = ops_out + dense[x];
                o.cells = r.cells;
            }
            ops_out += used;
        }
    }
    // jobs that did not run: capped (status 2) or without band cells (status 1)
    for (uint64_t j = 0; j < njobs; ++j) {
        const BandJobDev& d = dj[j];
        if (capped[j] || d.w <= 0 || d.rows <= 0) {
            out[j] = saber_hip_alignment{};
            out[j].status = capped[j] ? 2 : 1;
            out[j].op_off = ops_out;
        }
    }
    return SABER_HIP_OK;
}

}  // namespace

extern "C" {

int saber_hip_band_align(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                         const saber_hip_band_job* jobs, const int32_t* stop_at, uint64_t njobs,
                         const saber_hip_scheme* scheme, uint64_t max_dir_bytes, saber_hip_alignment* out,
                         uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity) {
    if (int r_ = injected(c)) return r_;
    return band_align_impl(c, queries, subjects, jobs, stop_at, njobs, scheme, max_dir_bytes, out, op_codes,
                           op_lens, op_capacity, 0, nullptr);
}

int saber_hip_band_align_time(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                              const saber_hip_band_job* jobs, const int32_t* stop_at, uint64_t njobs,
                              const saber_hip_scheme* scheme, uint64_t max_dir_bytes, saber_hip_alignment* out,
                              uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity, int iters,
                              double* kernel_ms) {
    if (iters < 1 || !kernel_ms) return fail(SABER_HIP_EINVAL, "bad arguments");
    return band_align_impl(c, queries, subjects, jobs, stop_at, njobs, scheme, max_dir_bytes, out, op_codes,
                           op_lens, op_capacity, iters, kernel_ms);
}

// Engine::traceback (engine.rs:1328-1393) and traceback_from_end
// (:1403-1446), batched: every job's next banded_align_until attempt of a
// round goes to one band_align call, the score passes of the jobs falling
// back to their end cell in one score_bands call. Per job the decisions are
// the reference's, in the reference's order.
int saber_hip_traceback_ex(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                           const saber_hip_traceback_job* jobs, uint64_t njobs, const saber_hip_scheme* scheme,
                           int64_t band_pad, int exhaustive, uint64_t max_dir_bytes, saber_hip_alignment* out,
                           uint8_t* exact, uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity,
                           saber_hip_traceback_counters* counters, saber_hip_traceback_counters* job_counters) {
    if (int r_ = injected(c)) return r_;
    if (counters) *counters = saber_hip_traceback_counters{};
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (njobs == 0) return SABER_HIP_OK;
    if (!jobs || !out || !exact) return fail(SABER_HIP_EINVAL, "null argument");
    if (njobs >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "too many jobs");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    if (job_counters)
        for (uint64_t k = 0; k < njobs; ++k) job_counters[k] = saber_hip_traceback_counters{};
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "#   traceback %s %.2f ms\n", what, std::chrono::duration<double, std::milli>(t1 - tp).count());
        tp = t1;
    };
    std::vector<uint64_t> slen(njobs);   // each job's subject length
    uint64_t gen = 0;                    // the database this call reads (subjects == NULL)
    {
        std::lock_guard<std::mutex> g(c->mu);
        if (!subjects && !c->db.p) return fail(SABER_HIP_ESTATE, "no database uploaded");
        const uint64_t* so = subjects ? subjects->offsets : c->db_off.data();
        const uint32_t ns = subjects ? subjects->count : (uint32_t)c->db_off.size() - 1;
        for (uint64_t k = 0; k < njobs; ++k) {
            if (jobs[k].subject >= ns) return fail(SABER_HIP_EINVAL, "job subject out of range");
            slen[k] = so[jobs[k].subject + 1] - so[jobs[k].subject];
        }
        gen = c->db_gen;
    }
    // A database upload from another thread between two rounds would mix
    // two databases in one result: every round checks the generation.
    auto same_db = [&]() -> int {
        if (subjects) return SABER_HIP_OK;
        std::lock_guard<std::mutex> g(c->mu);
        if (c->db_gen != gen) return fail(SABER_HIP_ESTATE, "the database was replaced during the traceback");
        return SABER_HIP_OK;
    };
    // per-job counters beside the batch sums
    auto jc = [&](uint64_t k) -> saber_hip_traceback_counters* { return job_counters ? &job_counters[k] : nullptr; };
    struct Tb {
        int64_t m, n, pad;
        int64_t base_lo, base_hi;
        bool base_full, done;
        int phase;          // 0 region band, 1 around the end cell
        bool need_score;    // the end cell is not known yet
        int64_t scope_lo, scope_hi, d_end;
        int32_t target;
        int64_t cur_lo, cur_hi;
        bool covers;
        bool have_best;
        saber_hip_alignment best;
        std::vector<uint8_t> bc;
        std::vector<uint32_t> bl;
    };
    std::vector<Tb> T(njobs);
    std::vector<uint64_t> op_off(njobs);
    uint64_t ops_total = 0;
    const int64_t pad0 = std::max<int64_t>(band_pad, 4);
    for (uint64_t k = 0; k < njobs; ++k) {
        const saber_hip_traceback_job& J = jobs[k];
        if (J.query >= queries->count) return fail(SABER_HIP_EINVAL, "job query out of range");
        if (J.win_hi < J.win_lo || J.win_hi > slen[k]) return fail(SABER_HIP_EINVAL, "job window outside its subject");
        Tb& t = T[k];
        t.m = (int64_t)(queries->offsets[J.query + 1] - queries->offsets[J.query]);
        t.n = (int64_t)J.win_hi - J.win_lo;
        if (t.m >= (1ll << 31) || t.n >= (1ll << 31)) return fail(SABER_HIP_EINVAL, "sequence too long");
        t.pad = pad0;
        // Band::full(m, n) when exhaustive or best == 0, else the region's
        // diagonals in window coordinates; a base equal to the full band is
        // not widened
        const bool full = exhaustive || J.best == 0;
        t.base_lo = full ? -t.m : J.d_min - (int64_t)J.win_lo;
        t.base_hi = full ? t.n : J.d_max - (int64_t)J.win_lo;
        t.base_full = t.base_lo == -t.m && t.base_hi == t.n;
        t.done = false;
        t.phase = 0;
        t.need_score = false;
        t.have_best = false;
        t.scope_lo = J.has_limit ? J.limit_lo : -t.m;   // the scoring band, or Band::full
        t.scope_hi = J.has_limit ? J.limit_hi : t.n;
        op_off[k] = ops_total;
        ops_total += (uint64_t)(t.m + t.n);
        out[k] = saber_hip_alignment{};
        out[k].status = 1;
        out[k].op_off = op_off[k];
        exact[k] = 0;
    }
    if (ops_total > op_capacity || (ops_total > 0 && (!op_codes || !op_lens)))
        return fail(SABER_HIP_EINVAL, "operation buffers need the sum over jobs of query + window length entries");
    lap("set-up");
    saber_hip_traceback_counters cnt{};
    // a band in window diagonals as an int32 job band: values beyond the
    // rectangle move to its edge (one past it for an empty band), which
    // Band::clip maps to the same band
    auto job_band = [](int64_t lo, int64_t hi, int64_t m, int64_t n, saber_hip_band_job& o) {
        o.band_lo = (int32_t)std::min<int64_t>(std::max<int64_t>(lo, -m), n + 1);
        o.band_hi = (int32_t)std::min<int64_t>(std::max<int64_t>(hi, -m - 1), n);
    };
    // operations of finished jobs, packed in the order the jobs finish
    uint64_t ops_out = 0;
    auto finish = [&](uint64_t k, const saber_hip_alignment* a, const uint8_t* oc, const uint32_t* ol, bool ex) {
        Tb& t = T[k];
        t.done = true;
        if (!a) return;
        out[k] = *a;
        out[k].op_off = ops_out;
        if (a->nops) {
            std::memcpy(op_codes + ops_out, oc, a->nops);
            std::memcpy(op_lens + ops_out, ol, 4ull * a->nops);
            ops_out += a->nops;
        }
        exact[k] = ex ? 1 : 0;
    };
    // the cap branch: traceback_from_end found nothing, so the best
    // alignment so far, not exact
    auto give_up = [&](uint64_t k) {
        Tb& t = T[k];
        cnt.band_capped += 1;
        if (auto* p = jc(k)) p->band_capped += 1;
        if (t.have_best) finish(k, &t.best, t.bc.data(), t.bl.data(), false);
        else finish(k, nullptr, nullptr, nullptr, false);
    };
    std::vector<uint64_t> ids;
    std::vector<saber_hip_band_job> bj;
    std::vector<int32_t> stops;
    std::vector<saber_hip_alignment> res;
    HostBuf<uint8_t> rc;
    HostBuf<uint32_t> rl;
    std::vector<int32_t> esc;
    std::vector<uint32_t> eqe, ese;
    std::vector<uint64_t> ecells;
    AlignSeqs seqs;   // the residues, uploaded by the first round for all
    for (;;) {
        // score passes of the jobs that fell back to their end cell
        ids.clear();
        bj.clear();
        for (uint64_t k = 0; k < njobs; ++k)
            if (!T[k].done && T[k].need_score) {
                const saber_hip_traceback_job& J = jobs[k];
                saber_hip_band_job b{J.query, J.subject, J.win_lo, J.win_hi, 0, 0};
                job_band(T[k].scope_lo, T[k].scope_hi, T[k].m, T[k].n, b);
                ids.push_back(k);
                bj.push_back(b);
            }
        if (!ids.empty()) {
            esc.assign(ids.size(), 0);
            eqe.assign(ids.size(), 0);
            ese.assign(ids.size(), 0);
            ecells.assign(ids.size(), 0);
            if (int r = saber_hip_score_bands(c, queries, subjects, bj.data(), bj.size(), scheme, esc.data(),
                                              eqe.data(), ese.data(), ecells.data()))
                return r;
            for (size_t x = 0; x < ids.size(); ++x) {
                Tb& t = T[ids[x]];
                t.need_score = false;
                cnt.banded_cells += ecells[x];
                if (auto* p = jc(ids[x])) p->banded_cells += ecells[x];
                if (esc[x] <= 0) {
                    give_up(ids[x]);
                    continue;
                }
                cnt.end_anchored_tracebacks += 1;
                if (auto* p = jc(ids[x])) p->end_anchored_tracebacks += 1;
                t.phase = 1;
                t.target = std::max(esc[x], jobs[ids[x]].want);
                t.d_end = (int64_t)ese[x] - (int64_t)eqe[x];
                t.pad = pad0;
            }
        }
        // this round's attempts
        ids.clear();
        bj.clear();
        stops.clear();
        uint64_t round_ops = 0;
        for (uint64_t k = 0; k < njobs; ++k) {
            Tb& t = T[k];
            if (t.done) continue;
            const saber_hip_traceback_job& J = jobs[k];
            int64_t lo, hi;
            bool covers;
            int32_t stop;
            if (t.phase == 0) {
                lo = t.base_full ? t.base_lo : t.base_lo - t.pad;
                hi = t.base_full ? t.base_hi : t.base_hi + t.pad;
                covers = lo <= -(t.m - 1) && hi >= t.n - 1;
                if (J.has_limit) {
                    lo = std::max<int64_t>(lo, J.limit_lo);
                    hi = std::min<int64_t>(hi, J.limit_hi);
                    covers = covers || (lo <= J.limit_lo && hi >= J.limit_hi);
                }
                stop = J.want;
            } else {
                lo = std::max<int64_t>(t.d_end - t.pad, t.scope_lo);
                hi = std::min<int64_t>(t.d_end + t.pad, t.scope_hi);
                covers = lo <= t.scope_lo && hi >= t.scope_hi;
                stop = t.target;
            }
            t.cur_lo = lo;
            t.cur_hi = hi;
            t.covers = covers;
            saber_hip_band_job b{J.query, J.subject, J.win_lo, J.win_hi, 0, 0};
            job_band(lo, hi, t.m, t.n, b);
            ids.push_back(k);
            bj.push_back(b);
            stops.push_back(stop);
            round_ops += (uint64_t)(t.m + t.n);
        }
        if (ids.empty()) break;
        cnt.tracebacks += ids.size();
        if (job_counters)
            for (uint64_t k : ids) job_counters[k].tracebacks += 1;
        if (int r = same_db()) return r;
        res.assign(ids.size(), saber_hip_alignment{});
        uint8_t* rcp = rc.get(std::max<uint64_t>(round_ops, 1));
        uint32_t* rlp = rl.get(std::max<uint64_t>(round_ops, 1));
        lap("round set-up");
        if (int r = band_align_impl(c, queries, subjects, bj.data(), stops.data(), bj.size(), scheme, max_dir_bytes,
                                    res.data(), rcp, rlp, round_ops, 0, nullptr, &seqs))
            return r;
        lap("band_align");
        for (size_t x = 0; x < ids.size(); ++x) {
            const uint64_t k = ids[x];
            Tb& t = T[k];
            const saber_hip_alignment& a = res[x];
            const uint8_t* oc = rcp + a.op_off;
            const uint32_t* ol = rlp + a.op_off;
            if (t.phase == 0) {
                if (a.status == 0) {
                    cnt.banded_cells += a.cells;
                    if (auto* p = jc(k)) p->banded_cells += a.cells;
                    if (a.score >= jobs[k].want || t.covers) {
                        finish(k, &a, oc, ol, a.score >= jobs[k].want);
                        continue;
                    }
                    // kept in case the band later exceeds the cap
                    t.have_best = true;
                    t.best = a;
                    t.bc.assign(oc, oc + a.nops);
                    t.bl.assign(ol, ol + a.nops);
                } else if (a.status == 1) {
                    if (t.covers) {
                        finish(k, nullptr, nullptr, nullptr, false);
                        continue;
                    }
                } else {
                    t.need_score = true;   // over the cap: traceback_from_end
                    continue;
                }
            } else {
                if (a.status == 0) {
                    cnt.banded_cells += a.cells;
                    if (auto* p = jc(k)) p->banded_cells += a.cells;
                    if (a.score >= t.target || t.covers) {
                        finish(k, &a, oc, ol, a.score >= jobs[k].want);
                        continue;
                    }
                } else if (a.status == 1) {
                    if (t.covers) {
                        give_up(k);
                        continue;
                    }
                } else {
                    give_up(k);
                    continue;
                }
            }
            cnt.band_widenings += 1;
            if (auto* p = jc(k)) p->band_widenings += 1;
            t.pad *= 2;
        }
    }
    lap("last round's results");
    if (int r = same_db()) return r;
    {
        // the call's device buffers, released on the context's device
        std::lock_guard<std::mutex> g(c->mu);
        (void)hipSetDevice(c->device);
        seqs.q.release();
        seqs.sbuf.release();
        seqs.bias.release();
        seqs.work.release();
        seqs.dir.release();
        seqs.codes.release();
        seqs.lens.release();
        seqs.dcodes.release();
        seqs.dlens.release();
        seqs.dst_off.release();
        seqs.djobs.release();
        seqs.mat.release();
        seqs.stops.release();
        seqs.recs.release();
        seqs.dids.release();
        seqs.dorder.release();
        seqs.dwoff.release();
        seqs.ddoff.release();
        seqs.dooff.release();
    }
    if (counters) *counters = cnt;
    return SABER_HIP_OK;
}

int saber_hip_traceback(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                        const saber_hip_traceback_job* jobs, uint64_t njobs, const saber_hip_scheme* scheme,
                        int64_t band_pad, int exhaustive, uint64_t max_dir_bytes, saber_hip_alignment* out,
                        uint8_t* exact, uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity,
                        saber_hip_traceback_counters* counters) {
    return saber_hip_traceback_ex(c, queries, subjects, jobs, njobs, scheme, band_pad, exhaustive, max_dir_bytes,
                                  out, exact, op_codes, op_lens, op_capacity, counters, nullptr);
}

int saber_hip_score_windows(saber_hip_ctx* c, const uint8_t* query, uint32_t qlen,
                            const int8_t* bias, const uint8_t* const* windows,
                            const uint32_t* lens, uint32_t n, const saber_hip_scheme* scheme,
                            int32_t* out) {
    if (int r_ = injected(c)) return r_;
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (n == 0) return SABER_HIP_OK;
    if (!out || !windows || !lens) return fail(SABER_HIP_EINVAL, "null argument");
    if (qlen == 0) {
        std::fill(out, out + n, 0);
        return SABER_HIP_OK;
    }
    if (!query) return fail(SABER_HIP_EINVAL, "query is null");
    // Concatenate the windows into one subject set and score the whole
    // rectangles with the exhaustive kernel (one query x n windows).
    std::vector<uint64_t> off(n + 1, 0);
    for (uint32_t k = 0; k < n; ++k) {
        if (lens[k] && !windows[k]) return fail(SABER_HIP_EINVAL, "window pointer is null");
        off[k + 1] = off[k] + lens[k];
    }
    std::vector<uint8_t> res(off[n] ? off[n] : 1);
    for (uint32_t k = 0; k < n; ++k)
        if (lens[k]) std::memcpy(res.data() + off[k], windows[k], lens[k]);
    uint64_t qoff[2] = {0, qlen};
    saber_hip_seqset qs{query, qoff, 1, bias};
    saber_hip_seqset ss{res.data(), off.data(), n, nullptr};
    return saber_hip_score_full(c, &qs, &ss, scheme, out, nullptr, nullptr);
}

}  // extern "C"

struct saber_hip_xdrop_batch {
    saber_hip_ctx* ctx = nullptr;
    uint64_t nhits = 0;
    int32_t x = 0;
    DevBuf<uint8_t> q, s;
    DevBuf<int8_t> bias;
    DevBuf<uint64_t> qoff, soff;
    DevBuf<uint32_t> hits;
    DevBuf<short> table;
    DevBuf<uint32_t> counter;
    DevBuf<uint4> desc;
    DevBuf<uint4> res;               // per hit: score, q_lo, q_hi, steps
    const uint8_t* sres = nullptr;
    const uint64_t* soffp = nullptr;
    bool has_bias = false, wide = false, ran = false;
    int grid = 0;
    uint32_t chunk = 0;
};

namespace {

int xdrop_enqueue(saber_hip_xdrop_batch* b) {
    hipStream_t st = b->ctx->stream;
    HIP_TRY(fill_u32(b->counter.p, kXdropGroups * kXdropCounterStride, 0, st));
    XdropArgs a{};
    a.hits = b->hits.p;
    a.nhits = (uint32_t)b->nhits;
    a.chunk = b->chunk;
    a.qoff = b->qoff.p;
    a.qres = b->q.p + kGuard;
    a.qbias = b->has_bias ? b->bias.p + kGuard : nullptr;
    a.soff = b->soffp;
    a.sres = b->sres;
    a.table = b->table.p;
    a.x = b->x;
    a.next_chunk = b->counter.p;
    a.desc = b->desc.p;
    a.res = b->res.p;
    HIP_TRY(launch_xdrop(a, b->has_bias, b->wide, b->grid, st));
    return SABER_HIP_OK;
}

}  // namespace

extern "C" {

int saber_hip_xdrop_batch_create(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                 const saber_hip_seqset* subjects, const saber_hip_xdrop_hit* hits,
                                 uint64_t nhits, const saber_hip_scheme* scheme, int32_t x,
                                 saber_hip_xdrop_batch** out) {
    if (int r_ = injected(c)) return r_;
    if (!c || !out || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (nhits > 0 && !hits) return fail(SABER_HIP_EINVAL, "hits is null");
    if (nhits >= (1ull << 31)) return fail(SABER_HIP_EINVAL, "too many hits for one batch");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    for (int i = 0; i < 576; ++i)
        if (scheme->matrix[i] < -32767 || scheme->matrix[i] > 32767)
            return fail(SABER_HIP_EINVAL, "substitution score outside the int16 range");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t* soffs = subjects ? subjects->offsets : c->db_off.data();
    const uint32_t nsub = subjects ? subjects->count : (c->db_off.empty() ? 0 : (uint32_t)c->db_off.size() - 1);
    for (uint64_t k = 0; k < nhits; ++k) {
        const saber_hip_xdrop_hit& h = hits[k];
        if (h.query >= queries->count || h.subject >= nsub)
            return fail(SABER_HIP_EINVAL, "hit index out of range");
        if (h.qp >= queries->offsets[h.query + 1] - queries->offsets[h.query] ||
            h.sp >= soffs[h.subject + 1] - soffs[h.subject])
            return fail(SABER_HIP_EINVAL, "hit position outside its sequence");
    }
    auto b = std::make_unique<saber_hip_xdrop_batch>();
    b->ctx = c;
    b->nhits = nhits;
    b->x = x;
    b->has_bias = queries->bias != nullptr;
    std::vector<short> table(kXdropTable);
    for (int a = 0; a < 32; ++a)
        for (int cc = 0; cc < 256; ++cc) {
            const int32_t v = scheme->matrix[(a < 24 ? a : kCodeX) * 24 + (cc < 24 ? cc : kCodeX)];
            table[a * 256 + cc] = (short)v;
            if (v < -128 || v > 127) b->wide = true;
        }
    hipStream_t st = c->stream;
    const uint64_t nq = queries->count ? queries->offsets[queries->count] : 0;
    if (int r = upload_codes(b->q, queries->residues, nq, st)) return r;
    if (b->has_bias)
        if (int r = upload_guarded(b->bias, queries->bias, nq, st)) return r;
    if (int r = upload(b->qoff, queries->offsets, (uint64_t)queries->count + 1, st)) return r;
    if (subjects) {
        const uint64_t ns = subjects->count ? subjects->offsets[subjects->count] : 0;
        if (int r = upload_codes(b->s, subjects->residues, ns, st)) return r;
        if (int r = upload(b->soff, subjects->offsets, (uint64_t)subjects->count + 1, st)) return r;
        b->sres = b->s.p + kGuard;
        b->soffp = b->soff.p;
    } else {
        if (!c->db.p && nhits > 0) return fail(SABER_HIP_ESTATE, "no database uploaded");
        if (int r = upload(b->soff, c->db_off.data(), (uint64_t)c->db_off.size(), st)) return r;
        b->sres = c->db.p + kGuard;
        b->soffp = b->soff.p;
    }
    if (int r = upload(b->hits, reinterpret_cast<const uint32_t*>(hits), 4 * nhits, st)) return r;
    if (int r = upload(b->table, table.data(), table.size(), st)) return r;
    HIP_TRY(b->counter.alloc(kXdropGroups * kXdropCounterStride));
    const uint64_t n1 = nhits ? nhits : 1;
    HIP_TRY(b->desc.alloc(2 * n1));
    HIP_TRY(b->res.alloc(n1));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    // Persistent grid: as many blocks of 4 waves as are resident at once, never
    // more than the work; a group's hits are cut into chunks of about a
    // quarter of a wave's share (at least 64: one descriptor load).
    const uint64_t need = (nhits + 255) / 256;
    b->grid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)cus * xdrop_blocks_per_cu(b->has_bias, b->wide), need));
    const uint64_t waves = 4ull * (uint64_t)b->grid;
    b->chunk = (uint32_t)std::max<uint64_t>(64, nhits / (4 * waves));
    HIP_TRY(hipStreamSynchronize(st));
    *out = b.release();
    return SABER_HIP_OK;
}

int saber_hip_xdrop_batch_run(saber_hip_xdrop_batch* b) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (int r = xdrop_enqueue(b)) return r;
    b->ran = true;
    return SABER_HIP_OK;
}

int saber_hip_xdrop_batch_read(saber_hip_xdrop_batch* b, int32_t* score, uint32_t* q_lo,
                               uint32_t* q_hi, uint32_t* steps) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    if (!b->ran) return fail(SABER_HIP_ESTATE, "batch has not run");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const uint64_t n = b->nhits;
    if (n == 0) return SABER_HIP_OK;
    std::vector<uint4> r(n);
    HIP_TRY(hipMemcpyAsync(r.data(), b->res.p, sizeof(uint4) * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < n; ++i) {
        if (score) score[i] = (int32_t)r[i].x;
        if (q_lo) q_lo[i] = r[i].y;
        if (q_hi) q_hi[i] = r[i].z;
        if (steps) steps[i] = r[i].w;
    }
    return SABER_HIP_OK;
}

int saber_hip_xdrop_batch_time(saber_hip_xdrop_batch* b, int warmup, int iters, double* ms_run,
                               uint64_t* steps) {
    if (!b || iters < 1 || warmup < 0) return fail(SABER_HIP_EINVAL, "bad arguments");
    for (int i = 0; i < warmup; ++i)
        if (int r = saber_hip_xdrop_batch_run(b)) return r;
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it)
        if (int r = xdrop_enqueue(b)) return r;
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    b->ran = true;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_run) *ms_run = ms / iters;
    if (steps) {
        std::vector<uint4> r(b->nhits);
        if (b->nhits) HIP_TRY(hipMemcpy(r.data(), b->res.p, sizeof(uint4) * b->nhits, hipMemcpyDeviceToHost));
        uint64_t t = 0;
        for (const uint4& v : r) t += v.w;
        *steps = t;
    }
    return SABER_HIP_OK;
}

void saber_hip_xdrop_batch_destroy(saber_hip_xdrop_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int saber_hip_xdrop(saber_hip_ctx* c, const saber_hip_seqset* queries,
                    const saber_hip_seqset* subjects, const saber_hip_xdrop_hit* hits,
                    uint64_t nhits, const saber_hip_scheme* scheme, int32_t x, int32_t* score,
                    uint32_t* q_lo, uint32_t* q_hi, uint32_t* steps) {
    if (int r_ = injected(c)) return r_;
    if (nhits > 0 && !score) return fail(SABER_HIP_EINVAL, "score is null");
    saber_hip_xdrop_batch* b = nullptr;
    if (int r = saber_hip_xdrop_batch_create(c, queries, subjects, hits, nhits, scheme, x, &b)) return r;
    int r = saber_hip_xdrop_batch_run(b);
    if (r == SABER_HIP_OK) r = saber_hip_xdrop_batch_read(b, score, q_lo, q_hi, steps);
    saber_hip_xdrop_batch_destroy(b);
    return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Exhaustive full-matrix scoring: every query against every subject.
// ---------------------------------------------------------------------------

// One register class of the full-matrix kernels: its persistent kernel's
// segments (in the batch's segment buffer) and chunk count.
struct FullClass {
    int nt = 0;
    uint32_t seg_off = 0, nsegs = 0;
    uint64_t nchunks = 0;
    uint32_t grid = 0;
};

struct saber_hip_full_batch {
    saber_hip_ctx* ctx = nullptr;
    uint32_t nq = 0, ns = 0;
    bool has_bias = false;
    int32_t go = 0, ge = 0;
    DevBuf<uint8_t> q, s;
    DevBuf<int8_t> bias;
    DevBuf<uint64_t> qoff, soff;
    DevBuf<uint32_t> qorder, pairs;
    DevBuf<uint32_t> pair_table;
    DevBuf<int32_t> mat;
    bool ends = true;                      // end cells computed (else scores only)
    std::vector<FullClass> classes;
    std::vector<FullSeg> host_segs;
    DevBuf<FullSeg> segs;
    DevBuf<uint32_t> next;                 // claim counters, one per class (64 B apart)
    std::vector<uint64_t> hq, hs;          // host offsets
    uint64_t qlen_total = 0, slen_total = 0;
    const uint8_t* sres = nullptr;
    DevBuf<int32_t> score;
    DevBuf<uint32_t> q_end, s_end;
    DevBuf<uint32_t> flagged, nflagged;
    // exact int32 path: pairs with subjects longer than the widest shape (a
    // static list) and pairs the packed kernel flagged, one work slot per thread
    DevBuf<uint32_t> static_pairs, static_count;
    uint64_t nstatic = 0;
    DevBuf<int32_t> gwork;
    uint32_t gthreads = 0, gslot = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t passes = 0;                   // passes enqueued so far
    bool ran = false;
    uint64_t last_flagged = 0;
};

namespace {

constexpr uint32_t kCounterStride = 16;   // u32 between class counters

GenericArgs full_generic_args(saber_hip_full_batch* b) {
    GenericArgs a{};
    a.qres = b->q.p + kGuard;
    a.qbias = b->has_bias ? b->bias.p + kGuard : nullptr;
    a.sres = b->sres;
    a.mat = b->mat.p;
    a.go = b->go;
    a.ge = b->ge;
    a.work = b->gwork.p;
    return a;
}

// One pass, entirely on the device (so it can be captured as a graph): claim
// counters cleared, the class kernels (concurrent, on the context's work
// streams), then the int32 kernel over the static long-subject pairs and
// over the pairs the packed kernels flagged (count read on the device).
// ev_begin / ev_end bracket the packed kernels.
int full_enqueue(saber_hip_full_batch* b, hipEvent_t ev_begin, hipEvent_t ev_end) {
    saber_hip_ctx* c = b->ctx;
    hipStream_t st = c->stream;
    HIP_TRY(fill_u32(b->nflagged.p, 1, 0, st));
    HIP_TRY(fill_u32(b->next.p, kCounterStride * std::max<size_t>(1, b->classes.size()), 0, st));
    FullKernelArgs a{};
    a.pairs = b->pairs.p;
    a.qorder = b->qorder.p;
    a.qoff = b->qoff.p;
    a.qres = b->q.p + kGuard;
    a.qbias = b->has_bias ? b->bias.p + kGuard : nullptr;
    a.qres_len = b->qlen_total;
    a.soff = b->soff.p;
    a.sres = b->sres;
    a.pair_table = b->pair_table.p;
    a.go = b->go;
    a.ge = b->ge;
    a.nsub = b->ns;
    a.score = b->score.p;
    a.q_end = b->q_end.p;
    a.s_end = b->s_end.p;
    a.flagged = b->flagged.p;
    a.nflagged = b->nflagged.p;
    if (ev_begin) HIP_TRY(hipEventRecord(ev_begin, st));
    const size_t nc = b->classes.size();
    if (nc > 1) {
        HIP_TRY(hipEventRecord(c->fork, st));
        for (size_t k = 1; k < nc; ++k) HIP_TRY(hipStreamWaitEvent(c->work[k - 1], c->fork, 0));
    }
    for (size_t k = 0; k < nc; ++k) {
        const FullClass& fc = b->classes[k];
        a.segs = b->segs.p + fc.seg_off;
        a.nsegs = fc.nsegs;
        a.nchunks = (uint32_t)fc.nchunks;
        a.next_chunk = b->next.p + kCounterStride * k;
        HIP_TRY(launch_full_multi(b->has_bias, b->ends, fc.nt, a, fc.grid, k == 0 ? st : c->work[k - 1]));
    }
    if (nc > 1)
        for (size_t k = 1; k < nc; ++k) {
            HIP_TRY(hipEventRecord(c->join[k - 1], c->work[k - 1]));
            HIP_TRY(hipStreamWaitEvent(st, c->join[k - 1], 0));
        }
    if (ev_end) HIP_TRY(hipEventRecord(ev_end, st));
    const GenericArgs g = full_generic_args(b);
    if (b->nstatic)
        HIP_TRY(launch_full_int32(g, b->static_pairs.p, b->static_count.p, b->qoff.p, b->soff.p, b->ns, b->score.p,
                                  b->q_end.p, b->s_end.p, b->gthreads, b->gslot, st));
    HIP_TRY(launch_full_int32(g, b->flagged.p, b->nflagged.p, b->qoff.p, b->soff.p, b->ns, b->score.p, b->q_end.p,
                              b->s_end.p, b->gthreads, b->gslot, st));
    return SABER_HIP_OK;
}

// Enqueue one pass: the first directly, later ones from the graph captured
// on the second (as for band batches), unless graphs are switched off.
int full_enqueue_run(saber_hip_full_batch* b) {
    hipStream_t st = b->ctx->stream;
    if ((b->ctx->flags & SABER_HIP_NO_GRAPH) || b->passes++ == 0) return full_enqueue(b, nullptr, nullptr);
    if (!b->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int r = full_enqueue(b, nullptr, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (r != SABER_HIP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, st));
    return SABER_HIP_OK;
}

// Shape of a subject length: the narrowest shape with L*C >= length (ties:
// fewer lanes), or the context's tuned choice for that length's class; -1
// when no shape covers it.
int full_shape_for(const saber_hip_ctx* c, bool ends, uint32_t len) {
    int ns = 0;
    const FullShape* sh = full_shapes(ends, &ns);
    int best = -1;
    for (int v = 0; v < ns; ++v) {
        const int W = sh[v].L * sh[v].C;
        if ((int64_t)W < (int64_t)len) continue;
        if (best < 0 || W < sh[best].L * sh[best].C || (W == sh[best].L * sh[best].C && sh[v].L < sh[best].L))
            best = v;
    }
    if (best >= 0) {
        const auto& tuned = c->full_tuned[ends ? 1 : 0];
        const auto it = tuned.find(best);
        if (it != tuned.end()) best = it->second;
    }
    return best;
}

}  // namespace

extern "C" {

int saber_hip_full_batch_create(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                const saber_hip_seqset* subjects, const saber_hip_scheme* scheme, int ends,
                                saber_hip_full_batch** out) {
    if (int r_ = injected(c)) return r_;
    if (!c || !out || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    auto b = std::make_unique<saber_hip_full_batch>();
    b->ctx = c;
    b->nq = queries->count;
    b->has_bias = queries->bias != nullptr;
    b->go = scheme->gap_open;
    b->ge = scheme->gap_extend;
    b->ends = ends != 0;
    b->hq.assign(queries->offsets, queries->offsets + queries->count + 1);
    if (subjects)
        b->hs.assign(subjects->offsets, subjects->offsets + subjects->count + 1);
    else
        b->hs = c->db_off;
    if (b->hs.empty()) b->hs.push_back(0);
    b->ns = (uint32_t)b->hs.size() - 1;
    if ((uint64_t)b->nq * b->ns >= (1ull << 40)) return fail(SABER_HIP_EINVAL, "batch too large");
    if (b->ns >= kNone) return fail(SABER_HIP_EINVAL, "too many subjects");

    bool packed_ok = !(c->flags & SABER_HIP_FORCE_EXACT32);
    packed_ok = packed_ok && scheme->gap_open >= 0 && scheme->gap_open <= 1000 &&
                scheme->gap_extend >= 0 && scheme->gap_extend <= 1000;
    for (int i = 0; i < 576; ++i)
        if (scheme->matrix[i] < -128 || scheme->matrix[i] > 127) packed_ok = false;

    // Subjects by shape (longest first within a shape), paired two per lane;
    // subjects no shape covers go to the int32 kernel with every query.
    int nshape = 0;
    const FullShape* sh = full_shapes(b->ends, &nshape);
    std::vector<std::vector<uint32_t>> bucket(nshape);
    std::vector<uint32_t> len(b->ns), long_subjects;
    uint32_t max_s = 0;
    for (uint32_t si = 0; si < b->ns; ++si) {
        len[si] = (uint32_t)(b->hs[si + 1] - b->hs[si]);
        max_s = std::max(max_s, len[si]);
        const int v = packed_ok ? full_shape_for(c, b->ends, len[si]) : -1;
        if (v < 0)
            long_subjects.push_back(si);
        else
            bucket[v].push_back(si);
    }
    std::vector<uint32_t> pairs;
    std::vector<uint32_t> pair_off(nshape, 0), npairs(nshape, 0);
    for (int v = 0; v < nshape; ++v) {
        auto& s = bucket[v];
        std::stable_sort(s.begin(), s.end(), [&](uint32_t x, uint32_t y) { return len[x] > len[y]; });
        pair_off[v] = (uint32_t)(pairs.size() / 2);
        npairs[v] = (uint32_t)((s.size() + 1) / 2);
        for (size_t i = 0; i < s.size(); i += 2) {
            pairs.push_back(s[i]);
            pairs.push_back(i + 1 < s.size() ? s[i + 1] : kNone);
        }
    }
    // Queries longest first; segments per register class, costliest chunk
    // first (a chunk runs m + L - 1 steps of C columns).
    std::vector<uint32_t> qorder(b->nq);
    for (uint32_t k = 0; k < b->nq; ++k) qorder[k] = k;
    auto qlen = [&](uint32_t k) { return b->hq[k + 1] - b->hq[k]; };
    std::stable_sort(qorder.begin(), qorder.end(), [&](uint32_t x, uint32_t y) { return qlen(x) > qlen(y); });
    const uint64_t max_m = b->nq ? qlen(qorder[0]) : 0;
    std::vector<int> nts;
    for (int v = 0; v < nshape; ++v)
        if (npairs[v] && std::find(nts.begin(), nts.end(), sh[v].nt) == nts.end()) nts.push_back(sh[v].nt);
    std::sort(nts.begin(), nts.end(), std::greater<int>());
    if (b->nq > 0)
        for (int nt : nts) {
            std::vector<int> vs;
            for (int v = 0; v < nshape; ++v)
                if (npairs[v] && sh[v].nt == nt) vs.push_back(v);
            std::stable_sort(vs.begin(), vs.end(), [&](int x, int y) {
                return (max_m + sh[x].L - 1) * sh[x].C > (max_m + sh[y].L - 1) * sh[y].C;
            });
            if ((int)vs.size() > kMaxFullSegs) return fail(SABER_HIP_EINVAL, "too many shapes in a class");
            FullClass fc;
            fc.nt = nt;
            fc.seg_off = (uint32_t)b->host_segs.size();
            uint64_t start = 0;
            for (int v : vs) {
                FullSeg sg{};
                sg.start = (uint32_t)start;
                sg.shape = (uint32_t)v;
                sg.nch = (npairs[v] + 64 / sh[v].L - 1) / (64 / sh[v].L);
                sg.pair_off = pair_off[v];
                sg.npairs = npairs[v];
                b->host_segs.push_back(sg);
                start += (uint64_t)b->nq * sg.nch;
            }
            if (start >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "batch too large for one pass");
            fc.nsegs = (uint32_t)vs.size();
            fc.nchunks = start;
            // persistent blocks: as many as fit at once, fewer when the
            // chunks would not give every wave one
            const int per_cu = full_multi_blocks_per_cu(b->has_bias, b->ends, nt);
            fc.grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * per_cu,
                                                                         (start + nt / 64 - 1) / (nt / 64)));
            b->classes.push_back(fc);
        }
    if (b->classes.size() > (size_t)kStreams + 1) return fail(SABER_HIP_EINVAL, "too many register classes");
    // int32 pairs: long subjects with every query
    std::vector<uint32_t> spairs;
    for (uint32_t si : long_subjects)
        for (uint32_t qi = 0; qi < b->nq; ++qi) {
            spairs.push_back(qi);
            spairs.push_back(si);
        }
    b->nstatic = spairs.size() / 2;
    if (b->nstatic >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "batch too large");

    const std::vector<uint32_t> ptable = build_pair_table(scheme->matrix);
    hipStream_t st = c->stream;
    const uint64_t nqres = b->hq.back();
    if (int r = upload_codes(b->q, queries->residues, nqres, st)) return r;
    if (b->has_bias)
        if (int r = upload_guarded(b->bias, queries->bias, nqres, st)) return r;
    if (subjects) {
        if (int r = upload_codes(b->s, subjects->residues, b->hs.back(), st)) return r;
        b->sres = b->s.p + kGuard;
    } else {
        if (!c->db.p && b->ns > 0) return fail(SABER_HIP_ESTATE, "no database uploaded");
        b->sres = c->db.p + kGuard;
    }
    b->qlen_total = nqres;
    b->slen_total = b->hs.back();
    if (int r = upload(b->qoff, b->hq.data(), b->hq.size(), st)) return r;
    if (int r = upload(b->soff, b->hs.data(), b->hs.size(), st)) return r;
    if (int r = upload(b->qorder, qorder.data(), qorder.size(), st)) return r;
    if (int r = upload(b->pairs, pairs.data(), pairs.size(), st)) return r;
    if (int r = upload(b->segs, b->host_segs.data(), b->host_segs.size(), st)) return r;
    if (int r = upload(b->pair_table, ptable.data(), ptable.size(), st)) return r;
    if (int r = upload(b->mat, scheme->matrix, 576, st)) return r;
    if (b->nstatic) {
        if (int r = upload(b->static_pairs, spairs.data(), spairs.size(), st)) return r;
        const uint32_t cnt = (uint32_t)b->nstatic;
        if (int r = upload(b->static_count, &cnt, 1, st)) return r;
    }
    HIP_TRY(b->next.alloc(kCounterStride * std::max<size_t>(1, b->classes.size())));
    const uint64_t nout = std::max<uint64_t>(1, (uint64_t)b->nq * b->ns);
    HIP_TRY(b->score.alloc(nout));
    HIP_TRY(b->q_end.alloc(nout));
    HIP_TRY(b->s_end.alloc(nout));
    HIP_TRY(b->flagged.alloc(2 * nout));
    HIP_TRY(b->nflagged.alloc(1));
    // int32 work slots: two band rows (4 arrays) of the widest full band
    b->gslot = (uint32_t)(4 * (max_m + max_s + 1));
    b->gthreads = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(256, b->nstatic));
    HIP_TRY(b->gwork.alloc((uint64_t)b->gthreads * b->gslot));
    HIP_TRY(hipStreamSynchronize(st));
    *out = b.release();
    return SABER_HIP_OK;
}

int saber_hip_full_batch_run(saber_hip_full_batch* b) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (int r = full_enqueue_run(b)) return r;
    b->ran = true;
    return SABER_HIP_OK;
}

int saber_hip_full_batch_read(saber_hip_full_batch* b, int32_t* score, uint32_t* q_end,
                              uint32_t* s_end) {
    if (!b) return fail(SABER_HIP_EINVAL, "batch is null");
    if (!b->ran) return fail(SABER_HIP_ESTATE, "batch has not run");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const uint64_t n = (uint64_t)b->nq * b->ns;
    if (!b->ends && (q_end || s_end)) return fail(SABER_HIP_ESTATE, "batch computes scores only");
    if (n == 0) return SABER_HIP_OK;
    if (score) HIP_TRY(hipMemcpyAsync(score, b->score.p, 4 * n, hipMemcpyDeviceToHost, st));
    if (q_end) HIP_TRY(hipMemcpyAsync(q_end, b->q_end.p, 4 * n, hipMemcpyDeviceToHost, st));
    if (s_end) HIP_TRY(hipMemcpyAsync(s_end, b->s_end.p, 4 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

int saber_hip_full_batch_time(saber_hip_full_batch* b, int warmup, int iters, double* ms_run,
                              uint64_t* cells, uint64_t* rescued) {
    if (!b || iters < 1 || warmup < 0) return fail(SABER_HIP_EINVAL, "bad arguments");
    for (int i = 0; i < warmup; ++i)
        if (int r = saber_hip_full_batch_run(b)) return r;
    std::lock_guard<std::mutex> g(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it)
        if (int r = full_enqueue_run(b)) return r;
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    b->ran = true;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint32_t nf = 0;
    HIP_TRY(hipMemcpy(&nf, b->nflagged.p, 4, hipMemcpyDeviceToHost));
    b->last_flagged = nf;
    if (ms_run) *ms_run = ms / iters;
    if (cells) *cells = b->hq.back() * b->hs.back();
    if (rescued) *rescued = b->nstatic + b->last_flagged;
    return SABER_HIP_OK;
}

void saber_hip_full_batch_destroy(saber_hip_full_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    delete b;
}

int saber_hip_full_autotune(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                            const saber_hip_scheme* scheme, int ends, uint64_t target_cells,
                            saber_hip_tune_entry* report, int nreport, int* nclasses) {
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (int r = check_seqset(queries, "queries")) return r;
    if (subjects)
        if (int r = check_seqset(subjects, "subjects")) return r;
    if (nclasses) *nclasses = 0;
    if (target_cells == 0) target_cells = 50000000000ull;
    const int fl = ends ? 1 : 0;
    std::vector<uint64_t> soffs;
    std::map<int, int> saved;
    {
        std::lock_guard<std::mutex> g(c->mu);
        soffs = subjects ? std::vector<uint64_t>(subjects->offsets, subjects->offsets + subjects->count + 1) : c->db_off;
        saved = c->full_tuned[fl];
        c->full_tuned[fl].clear();   // descend from the modelled choice
    }
    if (soffs.size() < 2 || queries->count == 0) return SABER_HIP_OK;
    const uint32_t ns = (uint32_t)soffs.size() - 1;
    int nshape = 0;
    const FullShape* sh = full_shapes(ends != 0, &nshape);
    // Length classes: subjects by modelled shape.
    std::vector<uint32_t> lo(nshape, UINT32_MAX), hi(nshape, 0);
    std::vector<uint64_t> res(nshape, 0);
    for (uint32_t si = 0; si < ns; ++si) {
        const uint32_t len = (uint32_t)(soffs[si + 1] - soffs[si]);
        const int v = full_shape_for(c, ends != 0, len);
        if (v < 0) continue;
        lo[v] = std::min(lo[v], len);
        hi[v] = std::max(hi[v], len);
        res[v] += len;
    }
    // The sample: whole queries spread over the length order, up to
    // target_cells cells against every subject.
    const uint64_t stot = soffs.back() - soffs.front();
    std::vector<uint32_t> qord(queries->count);
    for (uint32_t k = 0; k < queries->count; ++k) qord[k] = k;
    auto qlen = [&](uint32_t k) { return queries->offsets[k + 1] - queries->offsets[k]; };
    std::stable_sort(qord.begin(), qord.end(), [&](uint32_t x, uint32_t y) { return qlen(x) > qlen(y); });
    uint64_t qtot = 0;
    for (uint32_t k = 0; k < queries->count; ++k) qtot += qlen(k);
    const uint64_t per_q = std::max<uint64_t>(1, qtot / queries->count) * std::max<uint64_t>(1, stot);
    const uint32_t nsq = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(queries->count, target_cells / per_q));
    std::vector<uint8_t> sq_res;
    std::vector<int8_t> sq_bias;
    std::vector<uint64_t> sq_off{0};
    for (uint32_t k = 0; k < nsq; ++k) {
        const uint32_t qi = qord[(uint64_t)k * queries->count / nsq];
        const uint64_t a = queries->offsets[qi], b = queries->offsets[qi + 1];
        sq_res.insert(sq_res.end(), queries->residues + a, queries->residues + b);
        if (queries->bias) sq_bias.insert(sq_bias.end(), queries->bias + a, queries->bias + b);
        sq_off.push_back(sq_res.size());
    }
    saber_hip_seqset sq{};
    sq.residues = sq_res.data();
    sq.offsets = sq_off.data();
    sq.count = nsq;
    sq.bias = queries->bias ? sq_bias.data() : nullptr;
    const uint64_t sample_cells = sq_res.size() * stot;
    auto trial = [&](Timing* out) -> int {
        saber_hip_full_batch* b = nullptr;
        if (int r = saber_hip_full_batch_create(c, &sq, subjects, scheme, ends, &b)) return r;
        bool warm = false;
        const int r = timed3([&](double* ms_out) -> int {
            const int rr = saber_hip_full_batch_time(b, warm ? 0 : 1, 3, ms_out, nullptr, nullptr);
            warm = true;
            return rr;
        }, out);
        saber_hip_full_batch_destroy(b);
        return r;
    };
    // Coordinate descent: classes in decreasing share of the residues; per
    // class every shape covering its longest subject up to twice its width,
    // kept when the sample's pass gets clearly faster (clearly_faster).
    std::vector<int> order;
    for (int v = 0; v < nshape; ++v)
        if (res[v] * 100 >= stot && res[v] > 0) order.push_back(v);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return res[x] > res[y]; });
    Timing best;
    int status = trial(&best);
    double best_ms = best.median;
    std::vector<int> chosen(nshape, -1);
    for (int v : order) {
        if (status != SABER_HIP_OK) break;
        chosen[v] = v;
        const int Wv = sh[v].L * sh[v].C;
        for (int u = 0; u < nshape; ++u) {
            const int Wu = sh[u].L * sh[u].C;
            if (u == v || (uint32_t)Wu < hi[v] || Wu > 2 * Wv) continue;
            {
                std::lock_guard<std::mutex> g(c->mu);
                c->full_tuned[fl][v] = u;
            }
            Timing t;
            status = trial(&t);
            if (status != SABER_HIP_OK) break;
            if (clearly_faster(t, best)) {
                best = t;
                best_ms = t.median;
                chosen[v] = u;
            }
        }
        std::lock_guard<std::mutex> g(c->mu);
        if (chosen[v] != v)
            c->full_tuned[fl][v] = chosen[v];
        else
            c->full_tuned[fl].erase(v);
    }
    if (status != SABER_HIP_OK) {
        std::lock_guard<std::mutex> g(c->mu);
        c->full_tuned[fl] = saved;
        return status;
    }
    int n = 0;
    for (int v : order) {
        if (report && n < nreport) {
            saber_hip_tune_entry& e = report[n];
            e = saber_hip_tune_entry{};
            e.w_lo = (int32_t)lo[v];
            e.w_hi = (int32_t)hi[v];
            e.L = sh[chosen[v]].L;
            e.C = sh[chosen[v]].C;
            e.K = 0;
            e.reserved = chosen[v] != v;
            e.cells = (uint64_t)sq_res.size() * res[v];
            e.gcups = best_ms > 0 ? (double)sample_cells / (best_ms * 1e-3) / 1e9 : 0.0;
        }
        ++n;
    }
    if (nclasses) *nclasses = n;
    return SABER_HIP_OK;
}

int saber_hip_score_full(saber_hip_ctx* c, const saber_hip_seqset* queries,
                         const saber_hip_seqset* subjects, const saber_hip_scheme* scheme,
                         int32_t* score, uint32_t* q_end, uint32_t* s_end) {
    if (int r_ = injected(c)) return r_;
    saber_hip_full_batch* b = nullptr;
    if (int r = saber_hip_full_batch_create(c, queries, subjects, scheme, q_end || s_end, &b)) return r;
    int r = saber_hip_full_batch_run(b);
    if (r == SABER_HIP_OK) r = saber_hip_full_batch_read(b, score, q_end, s_end);
    saber_hip_full_batch_destroy(b);
    return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Memory probe: a plain device copy of known size (16-byte vector loads and
// stores), used to calibrate HBM traffic counters and to report achievable
// copy bandwidth next to the kernels.
// ---------------------------------------------------------------------------

namespace {

__global__ __launch_bounds__(256) void copy_probe_kernel(const uint4* __restrict__ src,
                                                         uint4* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int saber_hip_copy_probe(saber_hip_ctx* c, uint64_t bytes, int iters, double* ms) {
    if (!c || bytes < 16 || iters < 1) return fail(SABER_HIP_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint64_t n = bytes / 16;
    DevBuf<uint4> a, b;
    HIP_TRY(a.alloc(n));
    HIP_TRY(b.alloc(n));
    HIP_TRY(hipMemsetAsync(a.p, 1, 16 * n, st));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)cus * 16, (n + 255) / 256);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    SABER_LAUNCH(copy_probe_kernel, dim3(grid), dim3(256), 0, st, a.p, b.p, n);
    HIP_TRY(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i)
        SABER_LAUNCH(copy_probe_kernel, dim3(grid), dim3(256), 0, st, a.p, b.p, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / iters;
    return SABER_HIP_OK;
}

// ---------------------------------------------------------------------------
// Learned region prefilter: the MLP of mlp.hip behind the C ABI.

namespace {

// Offsets of the parameter blocks in the flat fp32 array (and of their
// gradients in the gradient array, whose last entry holds the loss sum).
constexpr uint64_t kPW1 = 0, kPb1 = 64 * 256, kPW2 = kPb1 + 256, kPb2 = kPW2 + 256 * 256,
                   kPw3 = kPb2 + 256, kPb3 = kPw3 + 256, kNParams = kPb3 + 1;
static_assert(kNParams == SABER_HIP_MLP_NPARAMS, "parameter layout");

}  // namespace

struct saber_hip_mlp {
    saber_hip_ctx* ctx = nullptr;
    DevBuf<float> p, g, am, av;        // parameters, gradients (+ loss), Adam moments
    DevBuf<__bf16> W1T, W2P, W2Q;      // bf16 operand layouts of W1, W2
    uint64_t steps = 0;                // Adam steps taken
    uint64_t cap = 0;                  // rows the batch buffers hold
    DevBuf<__bf16> X, H1, H2, dH1, dH2;
    DevBuf<float> y, lab, dy;
    DevBuf<__bf16> data_x;             // resident data set
    DevBuf<float> data_y;
    uint64_t data_n = 0;
    int max_grid = 1;
    int cus = 1;
};

namespace {

MlpWeights mlp_weights(const saber_hip_mlp* m) {
    return MlpWeights{m->W1T.p, m->W2P.p, m->W2Q.p, m->p.p + kPb1, m->p.p + kPb2, m->p.p + kPw3, m->p.p + kPb3};
}

MlpGrads mlp_grads(const saber_hip_mlp* m) {
    return MlpGrads{m->g.p + kPW1, m->g.p + kPW2, m->g.p + kPb1, m->g.p + kPb2,
                    m->g.p + kPw3, m->g.p + kPb3, m->g.p + kNParams};
}

int mlp_grid(const saber_hip_mlp* m, uint64_t n) {
    const uint64_t tiles = (n + 15) / 16, blocks = (tiles + 3) / 4;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)m->max_grid));
}

int mlp_reserve(saber_hip_mlp* m, uint64_t n) {
    if (n <= m->cap) return SABER_HIP_OK;
    HIP_TRY(m->X.alloc(n * 64));
    HIP_TRY(m->H1.alloc(n * 256));
    HIP_TRY(m->H2.alloc(n * 256));
    HIP_TRY(m->dH1.alloc(n * 256));
    HIP_TRY(m->dH2.alloc(n * 256));
    HIP_TRY(m->y.alloc(n));
    HIP_TRY(m->lab.alloc(n));
    HIP_TRY(m->dy.alloc(n));
    m->cap = n;
    return SABER_HIP_OK;
}

// Gradient of the batch loss into m->g (zeroed first); X, labels on device.
int mlp_backward_dev(saber_hip_mlp* m, const __bf16* X, const float* labels, uint64_t n, float pos_weight) {
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(m->g.p), kNParams + 1, 0, st));
    const MlpGrads gr = mlp_grads(m);
    HIP_TRY(mlp_launch_backward_act(mlp_weights(m), X, labels, n, pos_weight, 1.0f / (float)n, gr, m->H1.p,
                                    m->H2.p, m->dH2.p, m->dH1.p, m->dy.p, mlp_grid(m, n), st));
    // dW2 = H1^T dH2, dW1 = X^T dH1 (split over rows: about 1024 blocks each)
    HIP_TRY(mlp_launch_gemm_tn(m->H1.p, 256, 256, m->dH2.p, 256, 256, gr.dW2, n, 64, st));
    HIP_TRY(mlp_launch_gemm_tn(X, 64, 64, m->dH1.p, 256, 256, gr.dW1, n, 256, st));
    HIP_TRY(mlp_launch_colsum(m->dH2.p, nullptr, n, gr.db2, st));
    HIP_TRY(mlp_launch_colsum(m->dH1.p, nullptr, n, gr.db1, st));
    HIP_TRY(mlp_launch_colsum(m->H2.p, m->dy.p, n, gr.dw3, st));
    return SABER_HIP_OK;
}

int mlp_adam(saber_hip_mlp* m, float lr) {
    hipStream_t st = m->ctx->stream;
    m->steps += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, (double)m->steps));
    const float c2 = (float)(1.0 - std::pow(b2, (double)m->steps));
    HIP_TRY(mlp_launch_adam(m->p.p, m->g.p, m->am.p, m->av.p, (uint32_t)kNParams, lr, (float)b1, (float)b2,
                            1e-8f, c1, c2, st));
    HIP_TRY(mlp_launch_pack(m->p.p + kPW1, m->p.p + kPW2, m->W1T.p, m->W2P.p, m->W2Q.p, st));
    return SABER_HIP_OK;
}

int mlp_read_loss(saber_hip_mlp* m, uint64_t n, float* loss) {
    hipStream_t st = m->ctx->stream;
    float sum = 0.f;
    HIP_TRY(hipMemcpyAsync(&sum, m->g.p + kNParams, sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (loss) *loss = n ? sum / (float)n : 0.f;
    return SABER_HIP_OK;
}

int mlp_set(saber_hip_mlp* m, const float* params) {
    hipStream_t st = m->ctx->stream;
    HIP_TRY(hipMemcpyAsync(m->p.p, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, st));
    HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(m->am.p), kNParams, 0, st));
    HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(m->av.p), kNParams, 0, st));
    HIP_TRY(mlp_launch_pack(m->p.p + kPW1, m->p.p + kPW2, m->W1T.p, m->W2P.p, m->W2Q.p, st));
    HIP_TRY(hipStreamSynchronize(st));
    m->steps = 0;
    return SABER_HIP_OK;
}

}  // namespace

int saber::mlp_forward_device(saber_hip_mlp* m, const __bf16* X, uint64_t n, float* logits, hipStream_t st) {
    if (!m) return fail(SABER_HIP_EINVAL, "null model");
    if (n == 0) return SABER_HIP_OK;
    HIP_TRY(mlp_launch_forward(mlp_weights(m), X, n, logits, m->cus, st));
    return SABER_HIP_OK;
}

saber_hip_ctx* saber::mlp_context(const saber_hip_mlp* m) { return m ? m->ctx : nullptr; }

extern "C" {

int saber_hip_mlp_create(saber_hip_ctx* c, const float* params, uint64_t seed, saber_hip_mlp** out) {
    if (!c || !out) return fail(SABER_HIP_EINVAL, "null argument");
    *out = nullptr;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    auto m = std::make_unique<saber_hip_mlp>();
    m->ctx = c;
    HIP_TRY(m->p.alloc(kNParams));
    HIP_TRY(m->g.alloc(kNParams + 1));
    HIP_TRY(m->am.alloc(kNParams));
    HIP_TRY(m->av.alloc(kNParams));
    HIP_TRY(m->W1T.alloc(256 * 64));
    HIP_TRY(m->W2P.alloc(256 * 256));
    HIP_TRY(m->W2Q.alloc(256 * 256));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    m->max_grid = std::max(1, cus * 8);
    m->cus = std::max(1, cus);
    std::vector<float> init;
    if (!params) {
        // He-uniform weights, zero biases (splitmix64 stream from seed)
        init.assign(kNParams, 0.f);
        uint64_t s = seed;
        auto uni = [&]() {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        const double a1 = std::sqrt(6.0 / 64), a2 = std::sqrt(6.0 / 256);
        for (uint64_t i = 0; i < 64 * 256; ++i) init[kPW1 + i] = (float)(a1 * uni());
        for (uint64_t i = 0; i < 256 * 256; ++i) init[kPW2 + i] = (float)(a2 * uni());
        for (uint64_t i = 0; i < 256; ++i) init[kPw3 + i] = (float)(a2 * uni());
        params = init.data();
    }
    if (int r = mlp_set(m.get(), params)) return r;
    *out = m.release();
    return SABER_HIP_OK;
}

int saber_hip_mlp_get_params(saber_hip_mlp* m, float* params) {
    if (!m || !params) return fail(SABER_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    HIP_TRY(hipMemcpyAsync(params, m->p.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return SABER_HIP_OK;
}

int saber_hip_mlp_set_params(saber_hip_mlp* m, const float* params) {
    if (!m || !params) return fail(SABER_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    return mlp_set(m, params);
}

int saber_hip_mlp_forward(saber_hip_mlp* m, const uint16_t* x, uint64_t n, float* logits) {
    if (!m || (n && (!x || !logits))) return fail(SABER_HIP_EINVAL, "null argument");
    if (n == 0) return SABER_HIP_OK;
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    HIP_TRY(hipMemcpyAsync(m->X.p, x, 2 * 64 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(mlp_launch_forward(mlp_weights(m), m->X.p, n, m->y.p, m->cus, st));
    HIP_TRY(hipMemcpyAsync(logits, m->y.p, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

int saber_hip_mlp_backward(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n,
                           float pos_weight, float* grads, float* loss) {
    if (!m || n == 0 || !x || !labels) return fail(SABER_HIP_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    HIP_TRY(hipMemcpyAsync(m->X.p, x, 2 * 64 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m->lab.p, labels, sizeof(float) * n, hipMemcpyHostToDevice, st));
    if (int r = mlp_backward_dev(m, m->X.p, m->lab.p, n, pos_weight)) return r;
    if (grads) HIP_TRY(hipMemcpyAsync(grads, m->g.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, st));
    return mlp_read_loss(m, n, loss);
}

int saber_hip_mlp_train_step(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n,
                             float pos_weight, float lr, float* loss) {
    if (!m || n == 0 || !x || !labels) return fail(SABER_HIP_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    HIP_TRY(hipMemcpyAsync(m->X.p, x, 2 * 64 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m->lab.p, labels, sizeof(float) * n, hipMemcpyHostToDevice, st));
    if (int r = mlp_backward_dev(m, m->X.p, m->lab.p, n, pos_weight)) return r;
    if (int r = mlp_adam(m, lr)) return r;
    return mlp_read_loss(m, n, loss);
}

int saber_hip_mlp_upload(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n) {
    if (!m || (n && !x)) return fail(SABER_HIP_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    HIP_TRY(m->data_x.alloc(n ? n * 64 : 1));
    HIP_TRY(m->data_y.alloc(n ? n : 1));
    if (n) HIP_TRY(hipMemcpyAsync(m->data_x.p, x, 2 * 64 * n, hipMemcpyHostToDevice, st));
    if (n && labels) HIP_TRY(hipMemcpyAsync(m->data_y.p, labels, sizeof(float) * n, hipMemcpyHostToDevice, st));
    if (n && !labels) HIP_TRY(fill_u32(reinterpret_cast<uint32_t*>(m->data_y.p), n, 0, st));
    HIP_TRY(hipStreamSynchronize(st));
    m->data_n = n;
    return SABER_HIP_OK;
}

int saber_hip_mlp_train_resident(saber_hip_mlp* m, uint64_t row0, uint64_t n, float pos_weight, float lr,
                                 float* loss) {
    if (!m || n == 0 || row0 + n > m->data_n) return fail(SABER_HIP_EINVAL, "rows outside the resident data");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    if (int r = mlp_backward_dev(m, m->data_x.p + 64 * row0, m->data_y.p + row0, n, pos_weight)) return r;
    if (int r = mlp_adam(m, lr)) return r;
    if (!loss) return SABER_HIP_OK;
    return mlp_read_loss(m, n, loss);
}

int saber_hip_mlp_forward_resident(saber_hip_mlp* m, uint64_t row0, uint64_t n, float* logits) {
    if (!m || row0 + n > m->data_n) return fail(SABER_HIP_EINVAL, "rows outside the resident data");
    if (n && !logits) return fail(SABER_HIP_EINVAL, "logits is null");
    if (n == 0) return SABER_HIP_OK;
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    HIP_TRY(mlp_launch_forward(mlp_weights(m), m->data_x.p + 64 * row0, n, m->y.p, m->cus, st));
    HIP_TRY(hipMemcpyAsync(logits, m->y.p, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

int saber_hip_mlp_time(saber_hip_mlp* m, uint64_t n, int warmup, int iters, double* fwd_ms, double* step_ms) {
    if (!m || n == 0 || n > m->data_n || iters < 1 || warmup < 0) return fail(SABER_HIP_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t st = m->ctx->stream;
    if (int r = mlp_reserve(m, n)) return r;
    // Training steps change the weights: time them on a copy of the state.
    std::vector<float> saved(kNParams), sm(kNParams), sv(kNParams);
    HIP_TRY(hipMemcpyAsync(saved.data(), m->p.p, 4 * kNParams, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(sm.data(), m->am.p, 4 * kNParams, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(sv.data(), m->av.p, 4 * kNParams, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint64_t saved_steps = m->steps;
    hipEvent_t ev[4];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    const MlpWeights w = mlp_weights(m);
    for (int i = 0; i < warmup; ++i) HIP_TRY(mlp_launch_forward(w, m->data_x.p, n, m->y.p, m->cus, st));
    HIP_TRY(hipEventRecord(ev[0], st));
    for (int i = 0; i < iters; ++i) HIP_TRY(mlp_launch_forward(w, m->data_x.p, n, m->y.p, m->cus, st));
    HIP_TRY(hipEventRecord(ev[1], st));
    for (int i = 0; i < iters + warmup; ++i) {
        if (i == warmup) HIP_TRY(hipEventRecord(ev[2], st));
        if (int r = mlp_backward_dev(m, m->data_x.p, m->data_y.p, n, 1.0f)) return r;
        if (int r = mlp_adam(m, 0.0f)) return r;
    }
    HIP_TRY(hipEventRecord(ev[3], st));
    HIP_TRY(hipEventSynchronize(ev[3]));
    float t1 = 0.f, t2 = 0.f;
    HIP_TRY(hipEventElapsedTime(&t1, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&t2, ev[2], ev[3]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (fwd_ms) *fwd_ms = t1 / iters;
    if (step_ms) *step_ms = t2 / iters;
    HIP_TRY(hipMemcpyAsync(m->p.p, saved.data(), 4 * kNParams, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m->am.p, sm.data(), 4 * kNParams, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m->av.p, sv.data(), 4 * kNParams, hipMemcpyHostToDevice, st));
    HIP_TRY(mlp_launch_pack(m->p.p + kPW1, m->p.p + kPW2, m->W1T.p, m->W2P.p, m->W2Q.p, st));
    HIP_TRY(hipStreamSynchronize(st));
    m->steps = saved_steps;
    return SABER_HIP_OK;
}

void saber_hip_mlp_destroy(saber_hip_mlp* m) {
    if (!m) return;
    {
        std::lock_guard<std::mutex> g(m->ctx->mu);
        (void)hipSetDevice(m->ctx->device);
        (void)hipStreamSynchronize(m->ctx->stream);
    }
    delete m;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Two-stage region scoring: the banded part of Engine::score_windows
// (src/search/engine.rs:1021-1099) for the regions of many queries at once.

extern "C" int saber_hip_score_regions(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                       const saber_hip_seqset* subjects, const uint32_t* qidx,
                                       const saber_hip_region* regs, const saber_hip_plan* plans, uint64_t n,
                                       const saber_hip_scheme* scheme, int64_t narrow_band, int64_t full_band,
                                       const double* thr, int32_t* out, uint8_t* rescored) {
    return saber::score_regions_impl(c, queries, subjects, qidx, regs, plans, n, scheme, narrow_band, full_band,
                                     thr, out, rescored, nullptr, nullptr);
}

int saber::score_regions_impl(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                              const uint32_t* qidx, const saber_hip_region* regs, const saber_hip_plan* plans,
                              uint64_t n, const saber_hip_scheme* scheme, int64_t narrow_band, int64_t full_band,
                              const double* thr, int32_t* out, uint8_t* rescored, uint64_t* banded_cells,
                              const uint8_t* skip) {
    if (banded_cells) *banded_cells = 0;
    if (!c || !queries || !scheme || (n && (!qidx || !regs || !plans || !out))) return fail(SABER_HIP_EINVAL, "null argument");
    if (int r = check_seqset(queries, "queries")) return r;
    const bool two_stage = narrow_band > 0 && narrow_band < full_band;
    if (two_stage && n && !thr) return fail(SABER_HIP_EINVAL, "thr is null");
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "#   score_regions %s %.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(t1 - tp).count());
        tp = t1;
    };
    // Jobs in region order, built in blocks of regions on the host threads:
    // narrow-band jobs of banded plans (two-stage), full-band jobs of the
    // rest. The first bad region (in region order) decides the error.
    constexpr uint64_t kBlk = 1 << 15;
    const uint64_t nblk = (n + kBlk - 1) / kBlk;
    std::vector<uint64_t> n_narrow(nblk + 1, 0), n_full(nblk + 1, 0);
    std::vector<uint64_t> bad_at(nblk, ~0ull);
    std::vector<int> bad_kind(nblk, 0);
    // 0 skipped, 1 narrow pass, 2 full pass (banded plan), 3 full pass (whole rectangle)
    auto kind_of = [&](uint64_t i) -> int {
        if (skip && skip[i]) return 0;
        if (!plans[i].banded) return 3;
        return two_stage ? 1 : 2;
    };
    parallel_for(nblk, [&](uint64_t blk) {
        const uint64_t i1 = std::min(n, (blk + 1) * kBlk);
        uint64_t a = 0, f = 0;
        for (uint64_t i = blk * kBlk; i < i1; ++i) {
            if (qidx[i] >= queries->count) {
                bad_at[blk] = i;
                bad_kind[blk] = 1;
                return;
            }
            if (rescored) rescored[i] = 0;
            const int k = kind_of(i);
            if (k == 0) {
                out[i] = 0;
                continue;
            }
            if (plans[i].hi < plans[i].lo || regs[i].subject > 0xFFFFFFFFull) {
                bad_at[blk] = i;
                bad_kind[blk] = 2;
                return;
            }
            (k == 1 ? a : f) += 1;
        }
        n_narrow[blk + 1] = a;
        n_full[blk + 1] = f;
    });
    for (uint64_t blk = 0; blk < nblk; ++blk)
        if (bad_kind[blk] == 1) return fail(SABER_HIP_EINVAL, "region query index out of range");
        else if (bad_kind[blk] == 2) return fail(SABER_HIP_EINVAL, "bad plan");
    for (uint64_t blk = 0; blk < nblk; ++blk) {
        n_narrow[blk + 1] += n_narrow[blk];
        n_full[blk + 1] += n_full[blk];
    }
    // the full pass: full-band jobs, then the narrow jobs rescored
    const uint64_t nn = n_narrow[nblk], nf0 = n_full[nblk];
    auto narrow = raw_array<saber_hip_band_job>(nn), full = raw_array<saber_hip_band_job>(nf0 + nn);
    auto narrow_of = raw_array<uint64_t>(nn), full_of = raw_array<uint64_t>(nf0 + nn);
    auto banded_job = raw_array<uint8_t>(nf0 + nn);  // full-pass job k comes from a banded plan
    uint64_t nfull = nf0;
    parallel_for(nblk, [&](uint64_t blk) {
        uint64_t a = n_narrow[blk], f = n_full[blk];
        for (uint64_t i = blk * kBlk, i1 = std::min(n, (blk + 1) * kBlk); i < i1; ++i) {
            const int k = kind_of(i);
            if (k == 0) continue;
            const saber_hip_region& r = regs[i];
            const saber_hip_plan& p = plans[i];
            const int64_t m = (int64_t)(queries->offsets[qidx[i] + 1] - queries->offsets[qidx[i]]);
            const int64_t nw = (int64_t)(p.hi - p.lo);
            saber_hip_band_job j{qidx[i], (uint32_t)r.subject, (uint32_t)p.lo, (uint32_t)p.hi, 0, 0};
            if (k == 1) {
                // narrow band around the region's own diagonals, clipped
                j.band_lo = (int32_t)std::max<int64_t>(r.d_min - narrow_band - (int64_t)p.lo, -(m - 1));
                j.band_hi = (int32_t)std::min<int64_t>(r.d_max + narrow_band - (int64_t)p.lo, nw - 1);
                narrow[a] = j;
                narrow_of[a++] = i;
                continue;
            }
            if (k == 2) {
                j.band_lo = (int32_t)p.band_lo;
                j.band_hi = (int32_t)p.band_hi;
            } else {
                // the whole window rectangle: the full band of query x window
                j.band_lo = (int32_t)(-(m - 1));
                j.band_hi = (int32_t)(nw - 1);
            }
            full[f] = j;
            full_of[f] = i;
            banded_job[f++] = k == 2 ? 1 : 0;
        }
    });
    lap("jobs");
    if (nn) {
        auto sc = raw_array<int32_t>(nn);
        auto cells = raw_array<uint64_t>(banded_cells ? nn : 0);
        if (int r = saber_hip_score_bands(c, queries, subjects, narrow.get(), nn, scheme, sc.get(), nullptr, nullptr,
                                          banded_cells ? cells.get() : nullptr))
            return r;
        lap("narrow pass");
        if (banded_cells)
            for (uint64_t k = 0; k < nn; ++k) *banded_cells += cells[k];
        // narrow scores at or above the region's cut are rescored over the
        // full band (appended to the full pass in narrow-job order)
        const uint64_t nb2 = (nn + kBlk - 1) / kBlk;
        std::vector<uint64_t> n_res(nb2 + 1, 0);
        parallel_for(nb2, [&](uint64_t blk) {
            uint64_t x = 0;
            for (uint64_t k = blk * kBlk, k1 = std::min(nn, (blk + 1) * kBlk); k < k1; ++k) {
                const uint64_t i = narrow_of[k];
                if ((double)sc[k] >= thr[i]) {
                    ++x;
                    if (rescored) rescored[i] = 1;
                } else {
                    out[i] = sc[k];
                }
            }
            n_res[blk + 1] = x;
        });
        for (uint64_t blk = 0; blk < nb2; ++blk) n_res[blk + 1] += n_res[blk];
        parallel_for(nb2, [&](uint64_t blk) {
            uint64_t f = nf0 + n_res[blk];
            for (uint64_t k = blk * kBlk, k1 = std::min(nn, (blk + 1) * kBlk); k < k1; ++k) {
                const uint64_t i = narrow_of[k];
                if ((double)sc[k] >= thr[i]) {
                    saber_hip_band_job j = narrow[k];
                    j.band_lo = (int32_t)plans[i].band_lo;
                    j.band_hi = (int32_t)plans[i].band_hi;
                    full[f] = j;
                    banded_job[f] = 1;
                    full_of[f++] = i;
                }
            }
        });
        nfull = nf0 + n_res[nb2];
        lap("threshold");
    }
    if (nfull) {
        auto sc = raw_array<int32_t>(nfull);
        auto cells = raw_array<uint64_t>(banded_cells ? nfull : 0);
        if (int r = saber_hip_score_bands(c, queries, subjects, full.get(), nfull, scheme, sc.get(), nullptr, nullptr,
                                          banded_cells ? cells.get() : nullptr))
            return r;
        if (banded_cells)
            for (uint64_t k = 0; k < nfull; ++k)
                if (banded_job[k]) *banded_cells += cells[k];
        parallel_for((nfull + kBlk - 1) / kBlk, [&](uint64_t blk) {
            for (uint64_t k = blk * kBlk, k1 = std::min(nfull, (blk + 1) * kBlk); k < k1; ++k) out[full_of[k]] = sc[k];
        });
        lap("full pass");
    }
    return SABER_HIP_OK;
}

// ---------------------------------------------------------------------------
// Prefilter features of candidate regions (features.hip).


static int region_features_impl(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                const saber_hip_seqset* subjects, const saber_hip_band_job* jobs, uint64_t n,
                                const saber_hip_xdrop_hit* hits, const int32_t* xscore, const uint32_t* xq_lo,
                                const uint32_t* xq_hi, uint32_t hits_per_region, int32_t hsp_threshold,
                                const float* mean, const float* inv_std, uint16_t* features, int iters,
                                double* kernel_ms) {
    if (!c || !queries || !subjects || (n && (!jobs || !features))) return fail(SABER_HIP_EINVAL, "null argument");
    if (hits_per_region > 64) return fail(SABER_HIP_EINVAL, "at most 64 hits per region");
    if (n && hits_per_region && (!hits || !xscore || !xq_lo || !xq_hi)) return fail(SABER_HIP_EINVAL, "null hit arrays");
    if ((mean == nullptr) != (inv_std == nullptr)) return fail(SABER_HIP_EINVAL, "mean and inv_std go together");
    if (int r = check_seqset(queries, "queries")) return r;
    if (int r = check_seqset(subjects, "subjects")) return r;
    for (uint64_t k = 0; k < n; ++k) {
        const saber_hip_band_job& J = jobs[k];
        if (J.query >= queries->count || J.subject >= subjects->count) return fail(SABER_HIP_EINVAL, "job index out of range");
        if (J.win_hi < J.win_lo || J.win_hi > subjects->offsets[J.subject + 1] - subjects->offsets[J.subject])
            return fail(SABER_HIP_EINVAL, "job window outside its subject");
    }
    if (n == 0) return SABER_HIP_OK;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint64_t nq = queries->offsets[queries->count], ns = subjects->offsets[subjects->count];
    const uint64_t nh = n * hits_per_region;
    DevBuf<saber_hip_band_job> dj;
    DevBuf<uint64_t> dqo, dso;
    DevBuf<uint8_t> dq, ds;
    DevBuf<uint32_t> dh, dlo, dhi;
    DevBuf<int32_t> dsc;
    DevBuf<__bf16> dout;
    DevBuf<float> dmean, dinv;
    if (mean) {
        if (int r = upload(dmean, mean, 64, st)) return r;
        if (int r = upload(dinv, inv_std, 64, st)) return r;
    }
    if (int r = upload(dj, jobs, n, st)) return r;
    if (int r = upload(dqo, queries->offsets, (uint64_t)queries->count + 1, st)) return r;
    if (int r = upload(dso, subjects->offsets, (uint64_t)subjects->count + 1, st)) return r;
    if (int r = upload_codes(dq, queries->residues, nq, st)) return r;
    if (int r = upload_codes(ds, subjects->residues, ns, st)) return r;
    if (int r = upload(dh, reinterpret_cast<const uint32_t*>(hits), 4 * nh, st)) return r;
    if (int r = upload(dsc, xscore, nh, st)) return r;
    if (int r = upload(dlo, xq_lo, nh, st)) return r;
    if (int r = upload(dhi, xq_hi, nh, st)) return r;
    HIP_TRY(dout.alloc(64 * n));
    DevBuf<float> dcomp;
    HIP_TRY(dcomp.alloc(20 * ((uint64_t)queries->count + 1)));
    FeatureArgs a{dj.p, n, dqo.p, dq.p + kGuard, dso.p, ds.p + kGuard, dh.p, dsc.p, dlo.p, dhi.p,
                  hits_per_region, hsp_threshold, mean ? dmean.p : nullptr, mean ? dinv.p : nullptr, dout.p,
                  dcomp.p};
    HIP_TRY(launch_region_features(a, queries->count, st));
    HIP_TRY(hipMemcpyAsync(features, dout.p, 2 * 64 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (iters > 0 && kernel_ms) {
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, st));
        for (int it = 0; it < iters; ++it) HIP_TRY(launch_region_features(a, queries->count, st));
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *kernel_ms = ms / iters;
    }
    return SABER_HIP_OK;
}

extern "C" int saber_hip_region_features(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                         const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                                         uint64_t n, const saber_hip_xdrop_hit* hits, const int32_t* xscore,
                                         const uint32_t* xq_lo, const uint32_t* xq_hi, uint32_t hits_per_region,
                                         int32_t hsp_threshold, const float* mean, const float* inv_std,
                                         uint16_t* features) {
    return region_features_impl(c, queries, subjects, jobs, n, hits, xscore, xq_lo, xq_hi, hits_per_region,
                                hsp_threshold, mean, inv_std, features, 0, nullptr);
}

extern "C" int saber_hip_region_features_time(saber_hip_ctx* c, const saber_hip_seqset* queries,
                                              const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                                              uint64_t n, const saber_hip_xdrop_hit* hits, const int32_t* xscore,
                                              const uint32_t* xq_lo, const uint32_t* xq_hi,
                                              uint32_t hits_per_region, int32_t hsp_threshold, const float* mean,
                                              const float* inv_std, uint16_t* features, int iters,
                                              double* kernel_ms) {
    if (iters < 1 || !kernel_ms) return fail(SABER_HIP_EINVAL, "iters must be >= 1 and kernel_ms set");
    return region_features_impl(c, queries, subjects, jobs, n, hits, xscore, xq_lo, xq_hi, hits_per_region,
                                hsp_threshold, mean, inv_std, features, iters, kernel_ms);
}
