// The scan schedule end to end (Engine::search_scan,
// src/search/engine/scan.rs:1261-1497) over the context's database.
//
// Device stages (one search call, all queries of the batch together):
//   seed tables     every query window of every shape whose care residues
//                   are seed letters and whose code is not over-frequent in
//                   the database, sorted by code (stable: batch order within
//                   a code), with an entry-offset table and a presence bitmap
//                   per code (build_batch, scan.rs:1500-1596)
//   scan            one wave per 64 database positions: seed code, presence,
//                   the wave's hits spread over its lanes (prefix sum), the
//                   32-residue window identity filter of every hit
//                   (filter_avx2, scan.rs:2242-2323), survivors appended as
//                   64-bit keys (query, subject, diagonal, query position)
//   sort            radix sort of the keys: every (query, subject) pair's
//                   survivors together, each diagonal's in query-position order
//   X-drop          every survivor extended (xdrop.hip; extensions are pure
//                   functions of the seed pair)
//   coverage        per diagonal, the visit order's coverage rule replayed
//                   over those results (visit_scan, scan.rs:2384-2431): the
//                   visit order of a (query, subject) pair is (query
//                   position, shape, subject position), and on one diagonal
//                   that is query-position order -- equal positions on one
//                   diagonal are the same seed pair, whose extensions are
//                   identical
//   regions         per (query, subject): HSPs in diagonal order joined,
//                   thresholded and cut to the best regions
//                   (finish_subject_into, scan.rs:2437-2474)
// Host stages: preparation (composition bias, statistics), the per-query cap
// under region_order, plan_windows, the two-stage thresholds, E-values and
// the reporting loop of finish (engine.rs:1127-1221) with make_hit
// (engine.rs:1447-1530); the scoring passes and tracebacks it asks for run on
// the device (score_regions_impl, saber_hip_traceback_ex).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/saber_hip.h"
#include "common.h"
#include "ctx.h"
#include "host_par.h"
#include "search_score.h"

using namespace saber;

namespace saber {

namespace {

constexpr uint64_t kBatchPad = 32;     // batch positions between queries (scan.rs:37)
constexpr uint8_t kQSent = 0xFE;       // query window byte of a non-seed residue or padding (scan.rs:44)
constexpr uint8_t kNonSeed = 0x80;     // database byte of a residue outside the seed alphabet
constexpr uint32_t kTile = 64;         // database positions per wave
constexpr uint32_t kMaxRegionsPerSubject = 8;
constexpr uint32_t kMaxWeight = 16;
constexpr uint64_t kMaxCodes = 1ull << 28;
constexpr uint32_t kOccCapFloor = 256;  // src/search/index.rs:321

struct ShapeDev {
    uint32_t w, span;
    uint8_t off[kMaxWeight];
};

// Bit layout of a survivor key: query | subject | diagonal + maxq | query
// position, most significant first; sorting the keys groups a (query,
// subject) pair's survivors by diagonal, each in query-position order.
struct KeyLayout {
    uint32_t bp, bd, bs, bq;
    int64_t dofs;
    __host__ __device__ uint64_t make(uint32_t q, uint32_t s, int64_t d, uint32_t qp) const {
        return ((uint64_t)q << (bp + bd + bs)) | ((uint64_t)s << (bp + bd)) | ((uint64_t)(d + dofs) << bp) | qp;
    }
    __host__ __device__ uint32_t qp(uint64_t k) const { return (uint32_t)(k & ((1ull << bp) - 1)); }
    __host__ __device__ int64_t d(uint64_t k) const { return (int64_t)((k >> bp) & ((1ull << bd) - 1)) - dofs; }
    __host__ __device__ uint32_t s(uint64_t k) const { return (uint32_t)((k >> (bp + bd)) & ((1ull << bs) - 1)); }
    __host__ __device__ uint32_t q(uint64_t k) const { return (uint32_t)(k >> (bp + bd + bs)); }
    __host__ __device__ uint64_t diag_seg(uint64_t k) const { return k >> bp; }
    __host__ __device__ uint64_t pair_seg(uint64_t k) const { return k >> (bp + bd); }
};

uint32_t bits_for(uint64_t x) {
    uint32_t b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

// Unaligned 32 bytes from global memory as eight dwords (nine aligned loads;
// the buffers are guarded, so the few bytes past the window are readable).
__device__ __forceinline__ void load32u(const uint8_t* p, uint32_t out[8]) {
    const uintptr_t a = (uintptr_t)p;
    const uint32_t* b = (const uint32_t*)(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3);
    uint32_t d[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = b[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

// 0x80 in every byte of x that is zero.
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// 0xFF in byte t of the dword for every bit t of m (4 bits).
__device__ __forceinline__ uint32_t expand4(uint32_t m) {
    return ((m & 1u) * 0xFFu) | ((m & 2u) * 0x7F80u) | ((m & 4u) * 0x3FC000u) | ((m & 8u) * 0x1FE00000u);
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(x, o, 64);
        if (lane >= o) x += v;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// ---- database side (cached per seed scheme) ----

struct LetterArgs {
    const uint8_t* db;
    uint64_t n;
    uint8_t let[32];   // alignment code -> seed letter << k, or kNonSeed
    uint8_t pk[32];    // alignment code -> packed byte (letter << k | rank), or kNonSeed
    uint8_t* dlet;
    uint8_t* dpk;
};

__global__ void db_letters_kernel(LetterArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t c = a.db[i] & 31;
    a.dlet[i] = a.let[c];
    a.dpk[i] = a.pk[c];
}

struct CountArgs {
    const uint2* tiles;   // (subject, first position) of every 64 positions
    uint64_t ntiles;
    const uint64_t* soff;
    const uint8_t* dlet;
    ShapeDev sh;
    uint32_t asize, kbits;
    uint32_t* counts;
    unsigned long long* total;
};

// Seed code of a database position from the letter bytes: digits are the
// seed letters at the care offsets (scan.rs:811-823); any non-seed residue
// makes the window invalid.
__device__ __forceinline__ bool db_code(const uint8_t* dl, const ShapeDev& sh, uint32_t asize, uint32_t kbits,
                                        uint32_t& code) {
    uint32_t c = 0;
    bool ok = true;
#pragma unroll
    for (uint32_t k = 0; k < kMaxWeight; ++k) {
        if (k < sh.w) {
            const uint32_t r = dl[sh.off[k]];
            ok = ok && !(r & kNonSeed);
            c = c * asize + (r >> kbits);
        }
    }
    code = c;
    return ok;
}

__global__ __launch_bounds__(256) void db_count_kernel(CountArgs a) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    uint64_t tot = 0;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < a.ntiles; t += nw) {
        const uint2 tl = a.tiles[t];
        const uint64_t sb = a.soff[tl.x];
        const uint32_t sl = (uint32_t)(a.soff[tl.x + 1] - sb);
        const uint32_t p = tl.y + lane;
        uint32_t code = 0;
        if (p + a.sh.span <= sl && db_code(a.dlet + sb + p, a.sh, a.asize, a.kbits, code)) {
            atomicAdd(&a.counts[code], 1u);
            ++tot;
        }
    }
    tot = wave_sum64(tot);
    if (lane == 0 && tot) atomicAdd(a.total, (unsigned long long)tot);
}

__global__ void freq_bits_kernel(const uint32_t* counts, uint64_t ncodes, uint64_t cap, uint64_t* bits) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 64 >= ncodes) return;
    uint64_t v = 0;
    for (uint32_t b = 0; b < 64 && w * 64 + b < ncodes; ++b)
        if (counts[w * 64 + b] > cap) v |= 1ull << b;
    bits[w] = v;
}

// ---- query side (per search) ----

struct QueryKeyArgs {
    const uint8_t* red;      // batch: seed letter of a seedable residue, else kQSent
    const uint32_t* qid;     // batch position -> query (kNone in padding)
    const uint64_t* qbase;   // first batch position of each query
    const uint64_t* qlen;
    uint64_t total;
    ShapeDev sh;
    uint32_t asize;
    uint32_t inv;            // ncodes: the key of an over-frequent seed; ncodes + 1: no seed
    const uint64_t* freq;    // over-frequent codes (null: none)
    uint32_t* keys;
    uint32_t* vals;
};

// The seed code of every batch position as a sort key. The counts of kept
// and over-frequent seeds are read off the sorted keys afterwards (the table
// offsets of codes ncodes and ncodes + 1), not counted here: one counter
// update per wave, from every wave of a short kernel, would serialise on the
// counter's address.
__global__ __launch_bounds__(256) void query_keys_kernel(QueryKeyArgs a) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = a.inv + 1;
    if (g < a.total) {
        const uint32_t q = a.qid[g];
        if (q != kNone) {
            const uint64_t p = g - a.qbase[q];
            if (p + a.sh.span <= a.qlen[q]) {
                uint32_t c = 0;
                bool ok = true;
#pragma unroll
                for (uint32_t k = 0; k < kMaxWeight; ++k) {
                    if (k < a.sh.w) {
                        const uint32_t r = a.red[g + a.sh.off[k]];
                        ok = ok && r < a.asize;
                        c = c * a.asize + (ok ? r : 0);
                    }
                }
                if (ok) key = (a.freq && ((a.freq[c >> 6] >> (c & 63)) & 1)) ? a.inv : c;
            }
        }
        a.keys[g] = key;
        a.vals[g] = (uint32_t)g;
    }
}

// seeds kept (offs[nc]) and over-frequent seeds skipped (offs[nc + 1] - offs[nc])
__global__ void table_counts_kernel(const uint32_t* offs, uint64_t nc, unsigned long long* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr[0] = offs[nc];
        ctr[1] = offs[nc + 1] - offs[nc];
    }
}

// offs[c] = first sorted entry with code >= c, for c in [0, last]
__global__ void table_offsets_kernel(const uint32_t* keys, uint32_t n, uint64_t last, uint32_t* offs) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > last) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    offs[c] = lo;
}

__global__ void table_bits_kernel(const uint32_t* offs, uint64_t ncodes, uint64_t* bits) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 64 >= ncodes) return;
    uint64_t v = 0;
    for (uint32_t b = 0; b < 64 && w * 64 + b < ncodes; ++b)
        if (offs[w * 64 + b + 1] > offs[w * 64 + b]) v |= 1ull << b;
    bits[w] = v;
}

// ---- the scan: seed lookups and window filter ----

struct ScanArgs {
    const uint2* tiles;
    uint64_t ntiles;
    const uint64_t* soff;
    const uint8_t* dlet;      // guarded database letter bytes
    const uint8_t* dpk;       // guarded database packed bytes
    ShapeDev sh;
    uint32_t asize, kbits;
    const uint64_t* bits;     // presence of each code in the batch's table
    const uint32_t* offs;     // entry range of each code
    const uint32_t* entries;  // batch positions, by code
    const uint8_t* win;       // guarded batch window bytes (packed, kQSent)
    const uint32_t* qid;
    const uint64_t* qbase;
    int32_t c_half;           // half the shortest shape span
    uint32_t rm;              // (0xFF << k) in every byte
    uint32_t min_ids, kx;
    KeyLayout kl;
    uint64_t* out;
    uint64_t cap;
    unsigned long long* nout;
    unsigned long long* ctr;  // [0] positions with table entries, [1] hits, [2] hits filtered
};

constexpr int kScanWaves = 4;

constexpr uint32_t kStage = 256;   // survivors staged per wave before one append

__device__ __forceinline__ void flush_staged(const ScanArgs& a, uint64_t* st, uint32_t n, uint32_t lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(a.nout, (unsigned long long)n);
    base = __shfl(base, 0, 64);
    for (uint32_t k = lane; k < n; k += 64)
        if (base + k < a.cap) a.out[base + k] = st[k];
}

__global__ __launch_bounds__(256) void scan_kernel(ScanArgs a) {
    __shared__ unsigned long long s_ctr[3];
    __shared__ uint32_t s_excl[kScanWaves][64];
    __shared__ uint32_t s_ea[kScanWaves][64];
    __shared__ uint32_t s_sv[kScanWaves][64][8];
    __shared__ uint32_t s_sx[kScanWaves][64][8];
    __shared__ uint64_t s_out[kScanWaves][kStage];
    uint32_t staged = 0;   // wave-uniform
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < 3) s_ctr[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    uint64_t n_cand = 0, n_hits = 0, n_pass = 0;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < a.ntiles; t += nw) {
        const uint2 tl = a.tiles[t];
        const uint32_t s = tl.x;
        const uint64_t sb = a.soff[s];
        const uint32_t sl = (uint32_t)(a.soff[s + 1] - sb);
        const uint32_t p = tl.y + lane;
        uint32_t nh = 0, ea = 0;
        uint32_t code = 0;
        if (p + a.sh.span <= sl && db_code(a.dlet + sb + p, a.sh, a.asize, a.kbits, code) &&
            ((a.bits[code >> 6] >> (code & 63)) & 1)) {
            ea = a.offs[code];
            nh = a.offs[code + 1] - ea;
        }
        if (nh) {
            // the subject window [ss, ss + 32) as seed letters and packed
            // residues; bytes outside the subject never match (0xFF)
            const int32_t ss = (int32_t)p + a.c_half - 16;
            uint32_t sv[8], sx[8];
            load32u(a.dlet + (int64_t)sb + ss, sv);
            load32u(a.dpk + (int64_t)sb + ss, sx);
            const int32_t lo_l = ss < 0 ? -ss : 0;
            const int32_t hi_l = min((int32_t)sl - ss, 32);
            const uint32_t lanes = (hi_l >= 32 ? 0xFFFFFFFFu : ((1u << hi_l) - 1u)) & ~((1u << lo_l) - 1u);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t out = ~expand4((lanes >> (4 * i)) & 15u);
                s_sv[wv][lane][i] = sv[i] | out;
                s_sx[wv][lane][i] = sx[i] | out;
            }
        }
        const uint32_t incl = wave_incl_sum(nh, lane);
        const uint32_t H = __shfl(incl, 63, 64);
        s_excl[wv][lane] = incl - nh;
        s_ea[wv][lane] = ea;
        n_cand += __popcll(__ballot(nh != 0));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        n_hits += H;
        for (uint32_t b0 = 0; b0 < H; b0 += 64) {
            const uint32_t h = b0 + lane;
            bool pass = false;
            uint64_t key = 0;
            if (h < H) {
                // the position holding hit h: the last lane whose first hit is <= h
                uint32_t j = 0;
#pragma unroll
                for (uint32_t st = 32; st; st >>= 1)
                    if (s_excl[wv][j + st] <= h) j += st;
                const uint32_t g = a.entries[s_ea[wv][j] + (h - s_excl[wv][j])];
                uint32_t qw[8];
                load32u(a.win + g + a.c_half - 16, qw);
                uint32_t ids = 0, ex = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t eq = zero_bytes((qw[i] & a.rm) ^ s_sv[wv][j][i]);
                    ids += __popc(eq);
                    ex += __popc(zero_bytes(qw[i] ^ s_sx[wv][j][i]) & eq);
                }
                pass = ids + a.kx * ex >= a.min_ids;
                if (pass) {
                    const uint32_t q = a.qid[g];
                    const uint32_t qp = (uint32_t)(g - a.qbase[q]);
                    const uint32_t sp = tl.y + j;
                    key = a.kl.make(q, s, (int64_t)sp - (int64_t)qp, qp);
                }
            }
            const uint64_t bal = __ballot(pass);
            const uint32_t cnt = (uint32_t)__popcll(bal);
            n_pass += cnt;
            if (cnt) {
                if (staged + cnt > kStage) {
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    flush_staged(a, s_out[wv], staged, lane);
                    staged = 0;
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
                if (pass) s_out[wv][staged + __popcll(bal & ((1ull << lane) - 1ull))] = key;
                staged += cnt;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    if (staged) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        flush_staged(a, s_out[wv], staged, lane);
    }
    // counters: summed in the workgroup first (one global update per
    // workgroup and counter instead of one per wave)
    if (lane == 0) {
        atomicAdd(&s_ctr[0], (unsigned long long)n_cand);
        atomicAdd(&s_ctr[1], (unsigned long long)n_hits);
        atomicAdd(&s_ctr[2], (unsigned long long)(n_hits - n_pass));
    }
    __syncthreads();
    if (threadIdx.x < 3 && s_ctr[threadIdx.x]) atomicAdd(&a.ctr[threadIdx.x], s_ctr[threadIdx.x]);
}

// ---- survivors -> X-drop descriptors ----

struct KeyDescArgs {
    const uint64_t* keys;
    uint64_t n;
    KeyLayout kl;
    const uint64_t* qoff;
    const uint64_t* soff;
    XdropDescDev* desc;
};

__global__ void key_desc_kernel(KeyDescArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t k = a.keys[i];
    const uint32_t q = a.kl.q(k), s = a.kl.s(k), qp = a.kl.qp(k);
    const uint32_t sp = (uint32_t)((int64_t)qp + a.kl.d(k));
    const uint64_t q0 = a.qoff[q], s0 = a.soff[s];
    const uint32_t m = (uint32_t)(a.qoff[q + 1] - q0), sl = (uint32_t)(a.soff[s + 1] - s0);
    XdropDescDev d;
    d.q_abs = q0 + qp;
    d.s_abs = s0 + sp;
    d.qp = qp;
    d.n_right = min(m - qp, sl - sp);
    d.n_left = min(qp, sp);
    d.pad = 0;
    a.desc[i] = d;
}

// ---- diagonal coverage (visit_scan) ----

struct CoverArgs {
    const uint64_t* keys;
    uint64_t n;
    KeyLayout kl;
    const uint4* xres;        // score, q_lo, q_hi, steps
    int32_t hsp_threshold;
    uint8_t* hsp;             // 1: extended and an HSP
    unsigned long long* ctr;  // [0] extensions, [1] avoided, [2] steps, [3] (query, subject) groups
};

// One thread per diagonal: from the diagonal's first survivor, each
// extension's q_hi covers the survivors before it on the diagonal, so the
// next extended survivor is found by binary search (query positions rise
// along a diagonal); the survivors jumped over are the avoided ones. The HSP
// flags are cleared before the launch.
__global__ __launch_bounds__(256) void coverage_kernel(CoverArgs a) {
    __shared__ unsigned long long s_ctr[4];
    if (threadIdx.x < 4) s_ctr[threadIdx.x] = 0;
    __syncthreads();
    uint64_t ext = 0, avoided = 0, steps = 0, groups = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const uint64_t k = a.keys[i];
        const uint64_t seg = a.kl.diag_seg(k);
        const uint64_t prev = i ? a.keys[i - 1] : ~0ull;
        if (i == 0 || a.kl.pair_seg(prev) != a.kl.pair_seg(k)) ++groups;
        if (i == 0 || a.kl.diag_seg(prev) != seg) {
            // end of the diagonal: galloping search
            uint64_t lo = i + 1, step = 1, hi = i + 1;
            while (hi < a.n && a.kl.diag_seg(a.keys[hi]) == seg) {
                lo = hi + 1;
                hi = hi + step;
                step <<= 1;
            }
            if (hi > a.n) hi = a.n;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (a.kl.diag_seg(a.keys[mid]) == seg) lo = mid + 1;
                else hi = mid;
            }
            const uint64_t end = lo;
            uint64_t j = i, seg_ext = 0;
            while (j < end) {
                const uint4 r = a.xres[j];
                ++ext;
                ++seg_ext;
                steps += r.w;
                if ((int32_t)r.x >= a.hsp_threshold) a.hsp[j] = 1;
                const uint32_t cover = r.z;   // q_hi of this extension
                // first survivor after j with query position >= cover
                uint64_t l2 = j + 1, h2 = end;
                while (l2 < h2) {
                    const uint64_t mid = (l2 + h2) >> 1;
                    if (a.kl.qp(a.keys[mid]) < cover) l2 = mid + 1;
                    else h2 = mid;
                }
                j = l2;
            }
            avoided += (end - i) - seg_ext;
        }
    }
    ext = wave_sum64(ext);
    avoided = wave_sum64(avoided);
    steps = wave_sum64(steps);
    groups = wave_sum64(groups);
    // summed per workgroup, then one global update per workgroup and counter
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&s_ctr[0], (unsigned long long)ext);
        atomicAdd(&s_ctr[1], (unsigned long long)avoided);
        atomicAdd(&s_ctr[2], (unsigned long long)steps);
        atomicAdd(&s_ctr[3], (unsigned long long)groups);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_ctr[threadIdx.x]) atomicAdd(&a.ctr[threadIdx.x], s_ctr[threadIdx.x]);
}

// ---- HSPs -> regions (finish_subject_into) ----

struct RegionArgs {
    const uint64_t* keys;
    const uint4* xres;
    const uint32_t* hidx;     // survivor index of each HSP, in key order
    uint64_t nh;
    KeyLayout kl;
    int64_t join;             // 2 * window_slack
    int32_t hsp_threshold;
    const int32_t* region_thr;
    uint32_t maxr;            // max_regions_per_subject (<= 8)
    saber_hip_region* out;    // slot i.. of the pair's first HSP; reserved = query
    uint8_t* keep;
    unsigned long long* ctr;  // [0] regions dropped as weak
};

struct RegAcc {
    int64_t d_min, d_max;
    int32_t best, sum;
    uint32_t n;
};

// Grid-stride over the HSPs; the weak-region count is summed per workgroup
// and added to the counter once per workgroup.
__global__ __launch_bounds__(256) void region_kernel(RegionArgs a) {
    __shared__ unsigned long long s_weak;
    if (threadIdx.x == 0) s_weak = 0;
    __syncthreads();
    uint64_t weak = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nh;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k0 = a.keys[a.hidx[i]];
        const uint64_t pair = a.kl.pair_seg(k0);
        if (i == 0 || a.kl.pair_seg(a.keys[a.hidx[i - 1]]) != pair) {
            const uint32_t q = a.kl.q(k0), s = a.kl.s(k0);
            const int32_t thr = a.region_thr[q];
            const bool filt = thr > a.hsp_threshold;
            RegAcc top[kMaxRegionsPerSubject];
            uint32_t ntop = 0;
            auto emit = [&](const RegAcc& r) {
                if (filt && !(r.best >= thr || (r.n >= 2 && r.sum >= thr))) {
                    ++weak;
                    return;
                }
                // insert under (best desc, d_min asc), keep maxr
                uint32_t pos = ntop;
                while (pos > 0 && (top[pos - 1].best < r.best ||
                                   (top[pos - 1].best == r.best && top[pos - 1].d_min > r.d_min)))
                    --pos;
                if (pos >= a.maxr) return;
                const uint32_t last = ntop < a.maxr ? ntop : a.maxr - 1;
                for (uint32_t x = last; x > pos; --x) top[x] = top[x - 1];
                top[pos] = r;
                if (ntop < a.maxr) ++ntop;
            };
            RegAcc cur{0, 0, 0, 0, 0};
            bool have = false;
            uint64_t j = i;
            for (; j < a.nh; ++j) {
                const uint64_t kj = a.keys[a.hidx[j]];
                if (a.kl.pair_seg(kj) != pair) break;
                const int64_t d = a.kl.d(kj);
                const int32_t sc = (int32_t)a.xres[a.hidx[j]].x;
                if (have && d - cur.d_max <= a.join) {
                    cur.d_max = d > cur.d_max ? d : cur.d_max;
                    cur.best = sc > cur.best ? sc : cur.best;
                    cur.n += 1;
                    const int64_t sum = (int64_t)cur.sum + sc;
                    cur.sum = sum > INT32_MAX ? INT32_MAX : (sum < INT32_MIN ? INT32_MIN : (int32_t)sum);
                } else {
                    if (have) emit(cur);
                    cur = RegAcc{d, d, sc, sc, 1};
                    have = true;
                }
            }
            if (have) emit(cur);
            for (uint64_t x = i; x < j; ++x) {
                const uint64_t r = x - i;
                if (r < ntop) {
                    saber_hip_region o;
                    o.subject = s;
                    o.d_min = top[r].d_min;
                    o.d_max = top[r].d_max;
                    o.best = top[r].best;
                    o.n_hsp = top[r].n;
                    o.hsp_sum = top[r].sum;
                    o.reserved = q;
                    a.out[x] = o;
                    a.keep[x] = 1;
                } else {
                    a.keep[x] = 0;
                }
            }
        }
    }
    weak = wave_sum64(weak);
    if ((threadIdx.x & 63) == 0 && weak) atomicAdd(&s_weak, (unsigned long long)weak);
    __syncthreads();
    if (threadIdx.x == 0 && s_weak) atomicAdd(&a.ctr[0], s_weak);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

struct ScanDbCache {
    uint64_t db_gen = 0;
    std::string key;
    DevBuf<uint8_t> dlet, dpk;             // guarded like the residues
    DevBuf<uint2> tiles;
    uint64_t ntiles = 0;
    DevBuf<uint64_t> freq;                 // over-frequent code bitmaps, all shapes
    std::vector<int64_t> freq_off;         // per shape: first word in freq, -1 = none
    std::vector<uint64_t> caps, totals;
};

void scan_cache_free(ScanDbCache* p) { delete p; }

}  // namespace saber

struct saber_hip_search {
    uint32_t nq = 0;
    // the capped regions of all queries with their plans (feature export)
    host_vec<saber_hip_region> flat;
    host_vec<saber_hip_plan> plans;
    host_vec<uint32_t> qidx;
    std::vector<uint64_t> qlen;
    host_vec<uint64_t> slen;
    std::vector<int32_t> rthr;
    bool kept_regions = false;   // keep_regions: the arrays above and the region records are filled
    uint64_t db_gen = 0;
    std::vector<std::vector<saber_hip_hit>> hits;
    std::vector<std::vector<uint8_t>> op_codes;
    std::vector<std::vector<uint32_t>> op_lens;
    std::vector<std::vector<saber_hip_search_region>> regions;
    saber_hip_search_counters ctr{};
    saber_hip_search_timing tm{};
    // what the feature export needs
    saber_hip_ctx* ctx = nullptr;
    std::vector<uint8_t> qres;
    std::vector<uint64_t> qoff;
};

namespace saber {
namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct ParsedShape {
    ShapeDev dev;
    uint64_t ncodes;
};

// SeedShape::parse (alphabet.rs:364-384)
int parse_shape(const char* pat, uint32_t asize, ParsedShape& out) {
    const size_t n = strnlen(pat, SABER_HIP_MAX_SHAPE_SPAN + 4);
    if (n == 0 || n > SABER_HIP_MAX_SHAPE_SPAN) return fail(SABER_HIP_EINVAL, "seed shape span must be 1..64");
    if (pat[0] != '1' || pat[n - 1] != '1') return fail(SABER_HIP_EINVAL, "seed shape must start and end with 1");
    ShapeDev d{};
    d.span = (uint32_t)n;
    for (size_t i = 0; i < n; ++i) {
        if (pat[i] != '0' && pat[i] != '1') return fail(SABER_HIP_EINVAL, "seed shape must consist of 0/1");
        if (pat[i] == '1') {
            if (d.w >= 13) return fail(SABER_HIP_EINVAL, "seed shape weight must be in [2, 13]");
            d.off[d.w++] = (uint8_t)i;
        }
    }
    if (d.w < 2) return fail(SABER_HIP_EINVAL, "seed shape weight must be in [2, 13]");
    uint64_t nc = 1;
    for (uint32_t k = 0; k < d.w; ++k) {
        nc *= asize;
        if (nc > kMaxCodes) return fail(SABER_HIP_EINVAL, "seed code space above 2^28 codes (device table limit)");
    }
    out.dev = d;
    out.ncodes = nc;
    return SABER_HIP_OK;
}

// WinEnc::new (scan.rs:356-394): seed letter << k | rank within the letter's
// group, below 0x80.
struct WinEnc {
    uint32_t k = 0;
    uint8_t code[32];
    uint8_t letter[32];
};

int make_winenc(const saber_hip_seed_scheme* sc, WinEnc& e) {
    const uint32_t size = sc->alphabet_size;
    if (size == 0 || size > 31) return fail(SABER_HIP_EINVAL, "seed alphabet size must be 1..31");
    uint32_t count[32] = {0}, rank[24] = {0};
    for (int c = 0; c < 24; ++c) {
        const uint8_t r = sc->alphabet_map[c];
        if (r != 255) {
            if (r >= size) return fail(SABER_HIP_EINVAL, "seed alphabet map entry >= alphabet size");
            rank[c] = count[r]++;
        }
    }
    uint32_t m = 1;
    for (uint32_t x : count) m = std::max(m, x);
    uint32_t k = 0;
    while ((1u << k) < m) ++k;
    if ((size << k) > 0x80)
        return fail(SABER_HIP_EINVAL, "seed alphabet whose window bytes cannot be packed (letters << rank bits > 0x80)");
    e.k = k;
    std::memset(e.code, kNonSeed, sizeof e.code);
    std::memset(e.letter, kNonSeed, sizeof e.letter);
    for (int c = 0; c < 24; ++c) {
        const uint8_t r = sc->alphabet_map[c];
        if (r != 255) {
            e.code[c] = (uint8_t)((r << k) | rank[c]);
            e.letter[c] = (uint8_t)(r << k);
        }
    }
    return SABER_HIP_OK;
}

uint64_t auto_occ_cap(uint64_t total, uint64_t ncodes, uint64_t db_residues, uint64_t ref_residues) {
    uint64_t floor = kOccCapFloor;
    if (ref_residues > 0 && db_residues > ref_residues)
        floor = (uint64_t)std::ceil((double)kOccCapFloor * (double)db_residues / (double)ref_residues);
    const uint64_t c = (uint64_t)std::ceil((double)total / (double)std::max<uint64_t>(ncodes, 1) * 64.0);
    return std::max(c, floor);
}

int grid_for(uint64_t n, uint32_t block) {
    return (int)std::max<uint64_t>(1, (n + block - 1) / block);
}

// The database side of the scan for one seed scheme: letter bytes, position
// tiles and the over-frequent codes of every shape (build_scan_db,
// scan.rs:1065-1240, masking off).
int build_scan_cache(saber_hip_ctx* c, const saber_hip_seed_scheme* seeds, const WinEnc& enc,
                     const std::vector<ParsedShape>& shapes, const saber_hip_search_params* P, const std::string& key) {
    hipStream_t st = c->stream;
    auto cache = std::make_unique<ScanDbCache>();
    cache->db_gen = c->db_gen;
    cache->key = key;
    const uint64_t nres = c->db_off.empty() ? 0 : c->db_off.back();
    const uint64_t glen = nres + kGuard + kGuardTail;
    HIP_TRY(cache->dlet.alloc(glen));
    HIP_TRY(cache->dpk.alloc(glen));
    HIP_TRY(hipMemsetAsync(cache->dlet.p, 0xFF, glen, st));
    HIP_TRY(hipMemsetAsync(cache->dpk.p, 0xFF, glen, st));
    if (nres) {
        LetterArgs la{};
        la.db = c->db.p + kGuard;
        la.n = nres;
        std::memcpy(la.let, enc.letter, 32);
        std::memcpy(la.pk, enc.code, 32);
        la.dlet = cache->dlet.p + kGuard;
        la.dpk = cache->dpk.p + kGuard;
        SABER_LAUNCH(db_letters_kernel, dim3(grid_for(nres, 256)), dim3(256), 0, st, la);
        HIP_TRY(hipGetLastError());
    }
    // tiles of 64 positions, subject by subject
    std::vector<uint2> tiles;
    const uint32_t ns = (uint32_t)c->db_off.size() - 1;
    tiles.reserve(nres / kTile + ns + 1);
    for (uint32_t s = 0; s < ns; ++s) {
        const uint64_t sl = c->db_off[s + 1] - c->db_off[s];
        for (uint64_t p = 0; p < sl; p += kTile) tiles.push_back(make_uint2(s, (uint32_t)p));
    }
    cache->ntiles = tiles.size();
    if (int r = upload(cache->tiles, tiles.data(), tiles.size(), st)) return r;
    // over-frequent codes per shape
    uint64_t maxcodes = 1, words_total = 0;
    for (auto& sh : shapes) maxcodes = std::max(maxcodes, sh.ncodes);
    DevBuf<uint32_t> counts;
    DevBuf<unsigned long long> total;
    HIP_TRY(counts.alloc(maxcodes));
    HIP_TRY(total.alloc(1));
    cache->freq_off.assign(shapes.size(), -1);
    std::vector<uint64_t> need_bits;
    const int grid = c->cus * 8;
    for (size_t si = 0; si < shapes.size(); ++si) {
        const ParsedShape& sh = shapes[si];
        HIP_TRY(fill_u32(counts.p, sh.ncodes, 0, st));
        HIP_TRY(hipMemsetAsync(total.p, 0, sizeof(unsigned long long), st));
        if (cache->ntiles) {
            CountArgs ca{};
            ca.tiles = cache->tiles.p;
            ca.ntiles = cache->ntiles;
            ca.soff = c->db_off_dev.p;
            ca.dlet = cache->dlet.p + kGuard;
            ca.sh = sh.dev;
            ca.asize = seeds->alphabet_size;
            ca.kbits = enc.k;
            ca.counts = counts.p;
            ca.total = total.p;
            SABER_LAUNCH(db_count_kernel, dim3(grid), dim3(256), 0, st, ca);
            HIP_TRY(hipGetLastError());
        }
        unsigned long long tot = 0;
        HIP_TRY(hipMemcpyAsync(&tot, total.p, sizeof tot, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const uint64_t cap = P->seed_max_occ > 0 ? P->seed_max_occ
                                                 : auto_occ_cap(tot, sh.ncodes, nres, P->ref_db_residues);
        cache->caps.push_back(cap);
        cache->totals.push_back(tot);
        if (tot > cap) {
            // bitmaps of all such shapes in one buffer, filled below
            cache->freq_off[si] = (int64_t)words_total;
            words_total += (sh.ncodes + 63) / 64;
            need_bits.push_back(si);
        }
    }
    if (words_total) {
        HIP_TRY(cache->freq.alloc(words_total));
        for (uint64_t si : need_bits) {
            const ParsedShape& sh = shapes[si];
            // recount (one counts buffer for all shapes)
            HIP_TRY(fill_u32(counts.p, sh.ncodes, 0, st));
            HIP_TRY(hipMemsetAsync(total.p, 0, sizeof(unsigned long long), st));
            CountArgs ca{};
            ca.tiles = cache->tiles.p;
            ca.ntiles = cache->ntiles;
            ca.soff = c->db_off_dev.p;
            ca.dlet = cache->dlet.p + kGuard;
            ca.sh = sh.dev;
            ca.asize = seeds->alphabet_size;
            ca.kbits = enc.k;
            ca.counts = counts.p;
            ca.total = total.p;
            SABER_LAUNCH(db_count_kernel, dim3(grid), dim3(256), 0, st, ca);
            HIP_TRY(hipGetLastError());
            const uint64_t words = (sh.ncodes + 63) / 64;
            SABER_LAUNCH(freq_bits_kernel, dim3(grid_for(words, 256)), dim3(256), 0, st, counts.p, sh.ncodes,
                               cache->caps[si], cache->freq.p + cache->freq_off[si]);
            HIP_TRY(hipGetLastError());
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    c->scan.reset(cache.release());
    return SABER_HIP_OK;
}

std::string cache_key(const saber_hip_seed_scheme* seeds, const saber_hip_search_params* P) {
    std::string k(reinterpret_cast<const char*>(seeds->alphabet_map), 24);
    k += std::to_string(seeds->alphabet_size) + "|";
    for (uint32_t i = 0; i < seeds->nshapes; ++i) {
        k += std::string(seeds->shapes[i], strnlen(seeds->shapes[i], SABER_HIP_MAX_SHAPE_SPAN + 4));
        k += ",";
    }
    k += "|" + std::to_string(P->seed_max_occ) + "|" + std::to_string(P->ref_db_residues);
    return k;
}

// region_order (engine.rs:493-495)
bool region_before(const saber_hip_region& a, const saber_hip_region& b) {
    if (a.best != b.best) return a.best > b.best;
    if (a.subject != b.subject) return a.subject < b.subject;
    if (a.d_min != b.d_min) return a.d_min < b.d_min;
    return a.d_max < b.d_max;
}

struct QPrep {
    std::vector<int8_t> bias;
    bool has_bias = false;
    saber_hip_query_stats qs{};
    int64_t floor = 0, region_floor = 0;
    int32_t region_thr = 0;
};

struct HitCand {
    double e;
    uint64_t ri;
};

// Prefilter features of every region (bf16 [n][64] on the device): the
// query compositions from the device query codes, the windows from the
// database.
int region_features_dev(saber_hip_ctx* c, hipStream_t st, const uint64_t* d_qoff, const uint8_t* d_qres, uint32_t nq,
                        const saber_hip_region* regs, uint64_t n, const saber_hip_plan* plans,
                        const uint32_t* qidx, const uint64_t* qlen, const uint64_t* slen, const int32_t* rthr,
                        const float* mean, const float* inv_std, DevBuf<__bf16>& out) {
    auto rec = raw_array<SearchFeatRegion>(n);
    parallel_for((n + 4095) / 4096, [&](uint64_t blk) {
        for (uint64_t i = blk * 4096; i < std::min<uint64_t>(n, (blk + 1) * 4096); ++i) {
            const saber_hip_region& r = regs[i];
            const saber_hip_plan& pl = plans[i];
            SearchFeatRegion& x = rec[i];
            const uint32_t q = qidx[i];
            x.s_off = c->db_off[r.subject] + pl.lo;
            x.m = (uint32_t)qlen[q];
            x.n = (uint32_t)(pl.hi - pl.lo);
            x.band_lo = pl.banded ? (int32_t)pl.band_lo : -(int32_t)x.m + 1;
            x.band_hi = pl.banded ? (int32_t)pl.band_hi : (int32_t)x.n - 1;
            x.best = r.best;
            x.hsp_sum = r.hsp_sum;
            x.n_hsp = r.n_hsp;
            x.query = q;
            x.d_span = (int32_t)(r.d_max - r.d_min);
            x.thr = (float)rthr[q];
            x.slen = (uint32_t)slen[i];
            x.banded = pl.banded ? 1u : 0u;
        }
    });
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    DevBuf<SearchFeatRegion> d_rec;
    DevBuf<float> d_mean, d_inv, d_qcomp;
    if (int r = upload(d_rec, rec.get(), n, st)) return r;
    if (trace) {
        HIP_TRY(hipStreamSynchronize(st));
        std::fprintf(stderr, "#   features: %llu records of %zu bytes uploaded in %.1f ms\n", (unsigned long long)n,
                     sizeof(SearchFeatRegion),
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (mean && inv_std) {
        if (int r = upload(d_mean, mean, 64, st)) return r;
        if (int r = upload(d_inv, inv_std, 64, st)) return r;
    }
    HIP_TRY(d_qcomp.alloc(20ull * std::max<uint32_t>(nq, 1)));
    HIP_TRY(out.reserve(64 * std::max<uint64_t>(n, 1)));
    SearchFeatureArgs fa{};
    fa.regions = d_rec.p;
    fa.n = n;
    fa.sres = c->db.p + kGuard;
    fa.mean = (mean && inv_std) ? d_mean.p : nullptr;
    fa.inv_std = (mean && inv_std) ? d_inv.p : nullptr;
    fa.out = out.p;
    fa.qcomp = d_qcomp.p;
    HIP_TRY(launch_search_features(fa, d_qoff, d_qres, nq, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

}  // namespace
}  // namespace saber

extern "C" {

void saber_hip_search_defaults(saber_hip_search_params* p) {
    if (!p) return;
    *p = saber_hip_search_params{};
    p->evalue = 10.0;
    p->max_target_seqs = 10;
    p->max_hsps = 1;
    p->hsp_threshold = 40;
    p->x_drop_ungapped = 20;
    p->hsp_floor_fraction = 0.65;
    p->hit_filter_min_ids = 27;
    p->hit_filter_exact_weight = 2;
    p->band_pad = 16;
    p->window_slack = 64;
    p->score_band = 48;
    p->score_prefilter_band = 16;
    p->score_prefilter_frac = 0.75;
    p->max_candidates = 5000;
    p->max_regions_per_subject = 4;
    p->comp_bias = 1;
    p->seed_max_occ = 0;
    p->traceback_max_bytes = 32ull << 20;
    p->ref_db_residues = 40000000ull;
    p->per_subject_evalues = 1;
}

void saber_hip_seed_scheme_default(saber_hip_seed_scheme* s) {
    if (!s) return;
    *s = saber_hip_seed_scheme{};
    // Murphy-10 groups "A KR EDNQ C G H ILVM FYW P ST" over the codes
    // ARNDCQEGHILKMFPSTWYV
    static const char* groups[10] = {"A", "KR", "EDNQ", "C", "G", "H", "ILVM", "FYW", "P", "ST"};
    static const char* letters = "ARNDCQEGHILKMFPSTWYVBZX*";
    std::memset(s->alphabet_map, 255, 24);
    for (int g = 0; g < 10; ++g)
        for (const char* p = groups[g]; *p; ++p) s->alphabet_map[std::strchr(letters, *p) - letters] = (uint8_t)g;
    s->alphabet_size = 10;
    static const char* ds16[16] = {"11110111",          "11100100111",        "110010101011",
                                   "11010001001011",    "10101100001101",     "110100100100011",
                                   "1010010100010011",  "1100101000001011",   "11100000100010101",
                                   "11000100010010011", "11010000001000111",  "110001001000010011",
                                   "1010001000100001011", "1100010100000010011", "1100100000101000011",
                                   "1101000010000001011"};
    s->nshapes = 16;
    for (int i = 0; i < 16; ++i) std::strncpy(s->shapes[i], ds16[i], SABER_HIP_MAX_SHAPE_SPAN + 3);
}

int saber_hip_search_run(saber_hip_ctx* c, const saber_hip_seqset* queries, const saber_hip_seed_scheme* seeds,
                         const saber_hip_scheme* scoring, const saber_hip_karlin* karlin,
                         const saber_hip_search_params* P, const saber_hip_prefilter* prefilter,
                         saber_hip_search** out) {
    const auto t_start = Clock::now();
    if (!out) return fail(SABER_HIP_EINVAL, "out is null");
    *out = nullptr;
    if (!c || !seeds || !scoring || !karlin || !P) return fail(SABER_HIP_EINVAL, "null argument");
    if (int r = check_seqset(queries, "queries")) return r;
    if (seeds->nshapes == 0 || seeds->nshapes > SABER_HIP_MAX_SHAPES)
        return fail(SABER_HIP_EINVAL, "seed scheme needs 1..32 shapes");
    if (P->max_regions_per_subject == 0 || P->max_regions_per_subject > kMaxRegionsPerSubject)
        return fail(SABER_HIP_EINVAL, "max_regions_per_subject must be 1..8");
    if (P->max_target_seqs == 0 || P->max_hsps == 0) return fail(SABER_HIP_EINVAL, "max_target_seqs and max_hsps must be >= 1");
    if (prefilter && !prefilter->model) return fail(SABER_HIP_EINVAL, "prefilter without a model");
    WinEnc enc;
    if (int r = make_winenc(seeds, enc)) return r;
    std::vector<ParsedShape> shapes(seeds->nshapes);
    uint32_t min_span = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < seeds->nshapes; ++i) {
        if (int r = parse_shape(seeds->shapes[i], seeds->alphabet_size, shapes[i])) return r;
        min_span = std::min(min_span, shapes[i].dev.span);
    }
    const int32_t c_half = (int32_t)(min_span / 2);

    auto res = std::make_unique<saber_hip_search>();
    const uint32_t nq = queries->count;
    res->nq = nq;
    res->ctx = c;
    res->hits.resize(nq);
    res->op_codes.resize(nq);
    res->op_lens.resize(nq);
    res->regions.resize(nq);
    const uint64_t qtotal = nq ? queries->offsets[nq] : 0;
    res->qres.assign(queries->residues, queries->residues + qtotal);
    res->qoff.assign(queries->offsets, queries->offsets + nq + 1);
    if (nq == 0) {
        *out = res.release();
        return SABER_HIP_OK;
    }
    for (uint64_t i = 0; i < qtotal; ++i)
        if (res->qres[i] >= 24) res->qres[i] = (uint8_t)kCodeX;
    saber_hip_search_counters& ctr = res->ctr;
    saber_hip_search_timing& tm = res->tm;
    ctr.queries = nq;
    ctr.query_residues = qtotal;

    std::lock_guard<std::mutex> lock(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    if (!c->db.p || c->db_off.size() < 2) return fail(SABER_HIP_ESTATE, "no database uploaded");
    const uint32_t ns = (uint32_t)c->db_off.size() - 1;
    const uint64_t db_len = c->db_off.back();
    hipStream_t st = c->stream;

    // ---- preparation (Engine::prepare, engine.rs:759-786) ----
    auto t0 = Clock::now();
    std::vector<QPrep> prep(nq);
    const double region_e = [&]() {
        double e = std::max(P->evalue, 10.0);
        if (P->ref_db_residues > 0 && db_len > P->ref_db_residues)
            e *= (double)db_len / (double)P->ref_db_residues;
        return e;
    }();
    uint64_t maxq = 1;
    bool any_bias = false;
    // per query, on the host threads: composition bias, statistics, floors
    std::atomic<int> prep_err{SABER_HIP_OK};
    parallel_for(nq, [&](uint64_t q) {
        const uint64_t m = queries->offsets[q + 1] - queries->offsets[q];
        QPrep& pp = prep[q];
        pp.bias.assign(m, 0);
        if (P->comp_bias && m)
            saber_hip_composition_bias(queries->residues + queries->offsets[q], m, scoring->matrix, pp.bias.data());
        for (int8_t b : pp.bias) pp.has_bias = pp.has_bias || b != 0;
        if (int r = saber_hip_stats_for_query(karlin, db_len, ns, P->per_subject_evalues ? 1 : 0,
                                              std::max<uint64_t>(m, 1), &pp.qs)) {
            prep_err = r;
            return;
        }
        if (m == 0) {
            pp.floor = pp.region_floor = std::numeric_limits<int64_t>::max();
        } else {
            pp.floor = saber_hip_stats_min_score(&pp.qs, P->evalue, NAN);
            pp.region_floor = region_e == P->evalue ? pp.floor : saber_hip_stats_min_score(&pp.qs, region_e, NAN);
        }
        // hsp_threshold_for (engine.rs:731-739)
        const double f = P->hsp_floor_fraction;
        if (f > 0.0 && pp.region_floor > 0) {
            const double v = std::min(std::ceil((double)pp.region_floor * f), (double)INT32_MAX);
            pp.region_thr = std::max(P->hsp_threshold, (int32_t)v);
        } else {
            pp.region_thr = P->hsp_threshold;
        }
    });
    if (prep_err != SABER_HIP_OK) return fail(prep_err, "query statistics");
    for (uint32_t q = 0; q < nq; ++q) {
        maxq = std::max<uint64_t>(maxq, queries->offsets[q + 1] - queries->offsets[q]);
        any_bias = any_bias || prep[q].has_bias;
    }
    std::vector<int8_t> bias_all(qtotal, 0);
    for (uint32_t q = 0; q < nq; ++q)
        std::copy(prep[q].bias.begin(), prep[q].bias.end(), bias_all.begin() + queries->offsets[q]);
    tm.prepare = ms_since(t0);

    // ---- database side ----
    t0 = Clock::now();
    const std::string ck = cache_key(seeds, P);
    if (!c->scan || c->scan->db_gen != c->db_gen || c->scan->key != ck)
        if (int r = build_scan_cache(c, seeds, enc, shapes, P, ck)) return r;
    ScanDbCache* cache = c->scan.get();

    // ---- batch layout (build_batch) ----
    std::vector<uint64_t> base(nq), qlen(nq);
    uint64_t cur = kBatchPad;
    for (uint32_t q = 0; q < nq; ++q) {
        base[q] = cur;
        qlen[q] = queries->offsets[q + 1] - queries->offsets[q];
        cur += qlen[q] + kBatchPad;
    }
    const uint64_t total = cur;
    if (total >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "query batch above 2^32 batch positions");
    std::vector<uint8_t> red(total + kGuardTail, kQSent), win(total + kGuardTail, kQSent);
    std::vector<uint32_t> qid(total, kNone);
    for (uint32_t q = 0; q < nq; ++q)
        for (uint64_t i = 0; i < qlen[q]; ++i) {
            const uint8_t cc = queries->residues[queries->offsets[q] + i];
            const uint8_t r = seeds->alphabet_map[std::min<uint8_t>(cc, 23)];
            qid[base[q] + i] = q;
            if (r != 255) {
                red[base[q] + i] = r;
                win[base[q] + i] = enc.code[std::min<uint8_t>(cc, 23)];
            }
        }
    DevBuf<uint8_t> d_red, d_win, d_q;
    DevBuf<int8_t> d_bias;
    DevBuf<uint32_t> d_qid;
    DevBuf<uint64_t> d_qbase, d_qlen, d_qoff;
    if (int r = upload(d_red, red.data(), red.size(), st)) return r;
    if (int r = upload(d_win, win.data(), win.size(), st)) return r;
    if (int r = upload(d_qid, qid.data(), qid.size(), st)) return r;
    if (int r = upload(d_qbase, base.data(), base.size(), st)) return r;
    if (int r = upload(d_qlen, qlen.data(), qlen.size(), st)) return r;
    if (int r = upload(d_qoff, queries->offsets, (uint64_t)nq + 1, st)) return r;
    if (int r = upload_codes(d_q, queries->residues, qtotal, st)) return r;
    if (any_bias)
        if (int r = upload_guarded(d_bias, bias_all.data(), qtotal, st)) return r;

    // ---- seed tables (device) ----
    struct Table {
        DevBuf<uint32_t> keys, vals, offs;
        DevBuf<uint64_t> bits;
        uint64_t n = 0;
    };
    std::vector<std::unique_ptr<Table>> tabs;
    DevBuf<unsigned long long> d_tctr;
    HIP_TRY(d_tctr.alloc(2 * shapes.size()));
    DevBuf<uint32_t> tmp_keys;
    HIP_TRY(tmp_keys.alloc(total));
    DevBuf<uint8_t> sort_tmp;
    for (size_t si = 0; si < shapes.size(); ++si) {
        auto T = std::make_unique<Table>();
        const ParsedShape& sh = shapes[si];
        HIP_TRY(T->keys.alloc(total));
        HIP_TRY(T->vals.alloc(total));
        QueryKeyArgs qa{};
        qa.red = d_red.p;
        qa.qid = d_qid.p;
        qa.qbase = d_qbase.p;
        qa.qlen = d_qlen.p;
        qa.total = total;
        qa.sh = sh.dev;
        qa.asize = seeds->alphabet_size;
        qa.inv = (uint32_t)sh.ncodes;
        qa.freq = cache->freq_off[si] >= 0 ? cache->freq.p + cache->freq_off[si] : nullptr;
        qa.keys = tmp_keys.p;
        qa.vals = T->vals.p;
        SABER_LAUNCH(query_keys_kernel, dim3(grid_for(total, 256)), dim3(256), 0, st, qa);
        HIP_TRY(hipGetLastError());
        // stable sort by code; the values (batch positions) keep batch order per code
        DevBuf<uint32_t> vals_out;
        HIP_TRY(vals_out.alloc(total));
        size_t tb = 0;
        const unsigned end_bit = bits_for(sh.ncodes + 1);
        HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb, tmp_keys.p, T->keys.p, T->vals.p, vals_out.p, (size_t)total,
                                          0, end_bit, st));
        HIP_TRY(sort_tmp.reserve(tb));
        HIP_TRY(rocprim::radix_sort_pairs(sort_tmp.p, tb, tmp_keys.p, T->keys.p, T->vals.p, vals_out.p,
                                          (size_t)total, 0, end_bit, st));
        std::swap(T->vals.p, vals_out.p);
        std::swap(T->vals.n, vals_out.n);
        // code offsets over all the batch's keys (those of over-frequent
        // seeds and of positions without a seed sort last)
        const uint64_t nc = sh.ncodes;
        HIP_TRY(T->offs.alloc(nc + 2));
        HIP_TRY(T->bits.alloc((nc + 63) / 64));
        SABER_LAUNCH(table_offsets_kernel, dim3(grid_for(nc + 2, 256)), dim3(256), 0, st, T->keys.p, (uint32_t)total,
                     nc + 1, T->offs.p);
        SABER_LAUNCH(table_bits_kernel, dim3(grid_for((nc + 63) / 64, 256)), dim3(256), 0, st, T->offs.p, nc,
                     T->bits.p);
        SABER_LAUNCH(table_counts_kernel, dim3(1), dim3(64), 0, st, T->offs.p, nc, d_tctr.p + 2 * si);
        HIP_TRY(hipGetLastError());
        tabs.push_back(std::move(T));
    }
    std::vector<unsigned long long> tctr(2 * shapes.size());
    HIP_TRY(hipMemcpyAsync(tctr.data(), d_tctr.p, sizeof(unsigned long long) * tctr.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t si = 0; si < shapes.size(); ++si) {
        Table& T = *tabs[si];
        T.n = tctr[2 * si];
        ctr.seed_positions += tctr[2 * si];
        ctr.seeds_frequent_skipped += tctr[2 * si + 1];
        T.keys.release();
    }
    tmp_keys.release();
    tm.seed_tables = ms_since(t0);

    // ---- the scan ----
    t0 = Clock::now();
    uint64_t maxs = 1;
    for (uint32_t s = 0; s < ns; ++s) maxs = std::max<uint64_t>(maxs, c->db_off[s + 1] - c->db_off[s]);
    KeyLayout kl{};
    kl.bp = std::max<uint32_t>(1, bits_for(maxq));
    kl.bd = bits_for(maxq + maxs + 1);
    kl.bs = std::max<uint32_t>(1, bits_for(ns));
    kl.bq = std::max<uint32_t>(1, bits_for(nq));
    kl.dofs = (int64_t)maxq;
    const uint32_t key_bits = kl.bp + kl.bd + kl.bs + kl.bq;
    if (key_bits > 64)
        return fail(SABER_HIP_EINVAL, "query batch and database too large for 64-bit survivor keys: split the query batch");
    uint64_t cap = std::max<uint64_t>(1ull << 20, db_len * shapes.size() / 8);
    DevBuf<uint64_t> surv;
    HIP_TRY(surv.alloc(cap));
    DevBuf<unsigned long long> d_nout, d_sctr;
    HIP_TRY(d_nout.alloc(1));
    HIP_TRY(d_sctr.alloc(3 * shapes.size()));
    HIP_TRY(hipMemsetAsync(d_nout.p, 0, sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(d_sctr.p, 0, sizeof(unsigned long long) * 3 * shapes.size(), st));
    uint32_t rm = 0;
    for (int b = 0; b < 4; ++b) rm |= ((0xFFu << enc.k) & 0xFFu) << (8 * b);
    uint64_t nsurv = 0;
    for (size_t si = 0; si < shapes.size(); ++si) {
        if (tabs[si]->n == 0 || cache->ntiles == 0) continue;
        for (;;) {
            ScanArgs sa{};
            sa.tiles = cache->tiles.p;
            sa.ntiles = cache->ntiles;
            sa.soff = c->db_off_dev.p;
            sa.dlet = cache->dlet.p + kGuard;
            sa.dpk = cache->dpk.p + kGuard;
            sa.sh = shapes[si].dev;
            sa.asize = seeds->alphabet_size;
            sa.kbits = enc.k;
            sa.bits = tabs[si]->bits.p;
            sa.offs = tabs[si]->offs.p;
            sa.entries = tabs[si]->vals.p;
            sa.win = d_win.p;
            sa.qid = d_qid.p;
            sa.qbase = d_qbase.p;
            sa.c_half = c_half;
            sa.rm = rm;
            sa.min_ids = P->hit_filter_min_ids;
            sa.kx = P->hit_filter_exact_weight;
            sa.kl = kl;
            sa.out = surv.p;
            sa.cap = cap;
            sa.nout = d_nout.p;
            sa.ctr = d_sctr.p + 3 * si;
            const uint64_t waves = std::min<uint64_t>(cache->ntiles, (uint64_t)c->cus * 32);
            SABER_LAUNCH(scan_kernel, dim3(grid_for(waves * 64, 256)), dim3(256), 0, st, sa);
            HIP_TRY(hipGetLastError());
            unsigned long long n = 0;
            HIP_TRY(hipMemcpyAsync(&n, d_nout.p, sizeof n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (n <= cap) {
                nsurv = n;
                break;
            }
            // too many survivors: a larger buffer, keeping the earlier shapes', and this shape again
            const uint64_t ncap = std::max<uint64_t>(2 * cap, (uint64_t)((double)n * shapes.size() / (si + 1) * 1.25));
            DevBuf<uint64_t> bigger;
            HIP_TRY(bigger.alloc(ncap));
            if (nsurv)
                HIP_TRY(hipMemcpyAsync(bigger.p, surv.p, sizeof(uint64_t) * nsurv, hipMemcpyDeviceToDevice, st));
            std::swap(surv.p, bigger.p);
            std::swap(surv.n, bigger.n);
            cap = ncap;
            const unsigned long long keep = nsurv;
            HIP_TRY(hipMemcpyAsync(d_nout.p, &keep, sizeof keep, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemsetAsync(d_sctr.p + 3 * si, 0, sizeof(unsigned long long) * 3, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
    }
    {
        std::vector<unsigned long long> sc(3 * shapes.size());
        HIP_TRY(hipMemcpyAsync(sc.data(), d_sctr.p, sizeof(unsigned long long) * sc.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (size_t si = 0; si < shapes.size(); ++si) {
            ctr.scan_candidates += sc[3 * si];
            ctr.hits += sc[3 * si + 1];
            ctr.hits_filtered += sc[3 * si + 2];
        }
    }
    for (auto& T : tabs) T.reset();
    tabs.clear();
    tm.scan = ms_since(t0);
    if (nsurv >= (1ull << 32)) return fail(SABER_HIP_EINVAL, "more than 2^32 window-filter survivors: split the query batch");

    // ---- sort ----
    t0 = Clock::now();
    DevBuf<uint64_t> keys;
    HIP_TRY(keys.alloc(std::max<uint64_t>(nsurv, 1)));
    if (nsurv) {
        size_t tb = 0;
        HIP_TRY(rocprim::radix_sort_keys(nullptr, tb, surv.p, keys.p, (size_t)nsurv, 0, key_bits, st));
        HIP_TRY(sort_tmp.reserve(tb));
        HIP_TRY(rocprim::radix_sort_keys(sort_tmp.p, tb, surv.p, keys.p, (size_t)nsurv, 0, key_bits, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    surv.release();
    tm.sort = ms_since(t0);

    // ---- X-drop of every survivor ----
    t0 = Clock::now();
    DevBuf<XdropDescDev> desc;
    DevBuf<uint4> xres;
    HIP_TRY(desc.alloc(std::max<uint64_t>(nsurv, 1)));
    HIP_TRY(xres.alloc(std::max<uint64_t>(nsurv, 1)));
    bool wide = false;
    std::vector<short> xt(kXdropTable);
    for (int qa = 0; qa < 32; ++qa)
        for (int sc = 0; sc < 256; ++sc) {
            const int32_t v = scoring->matrix[(qa < 24 ? qa : kCodeX) * 24 + (sc < 24 ? sc : kCodeX)];
            if (v < -32767 || v > 32767) return fail(SABER_HIP_EINVAL, "substitution score outside the int16 range");
            xt[qa * 256 + sc] = (short)v;
            if (v < -128 || v > 127) wide = true;
        }
    DevBuf<short> d_xt;
    DevBuf<uint32_t> xcounter;
    if (int r = upload(d_xt, xt.data(), xt.size(), st)) return r;
    HIP_TRY(xcounter.alloc(kXdropGroups * kXdropCounterStride));
    if (nsurv) {
        KeyDescArgs ka{};
        ka.keys = keys.p;
        ka.n = nsurv;
        ka.kl = kl;
        ka.qoff = d_qoff.p;
        ka.soff = c->db_off_dev.p;
        ka.desc = desc.p;
        SABER_LAUNCH(key_desc_kernel, dim3(grid_for(nsurv, 256)), dim3(256), 0, st, ka);
        HIP_TRY(hipGetLastError());
        HIP_TRY(fill_u32(xcounter.p, kXdropGroups * kXdropCounterStride, 0, st));
        XdropArgs xa{};
        xa.hits = nullptr;
        xa.nhits = (uint32_t)nsurv;
        xa.qoff = d_qoff.p;
        xa.qres = d_q.p + kGuard;
        xa.qbias = any_bias ? d_bias.p + kGuard : nullptr;
        xa.soff = c->db_off_dev.p;
        xa.sres = c->db.p + kGuard;
        xa.table = d_xt.p;
        xa.x = P->x_drop_ungapped;
        xa.next_chunk = xcounter.p;
        xa.desc = desc.p;
        xa.res = xres.p;
        const int grid = (int)std::max<uint64_t>(
            1, std::min<uint64_t>((uint64_t)c->cus * xdrop_blocks_per_cu(any_bias, wide), (nsurv + 255) / 256));
        xa.chunk = (uint32_t)std::max<uint64_t>(64, nsurv / (16ull * grid));
        HIP_TRY(launch_xdrop_described(xa, any_bias, wide, grid, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    desc.release();
    tm.xdrop = ms_since(t0);

    // ---- coverage, HSPs, regions ----
    t0 = Clock::now();
    static const bool trace = std::getenv("SABER_HIP_TRACE_HOST") != nullptr;
    auto tl = Clock::now();
    auto lap = [&](const char* what) {
        if (!trace) return;
        (void)hipStreamSynchronize(st);
        std::fprintf(stderr, "#   search %s %.1f ms\n", what, ms_since(tl));
        tl = Clock::now();
    };
    DevBuf<saber_hip_region> rsel;   // region_kernel's selected regions, kept on the device
    uint64_t nr = 0;
    {
        DevBuf<uint8_t> hspf;
        DevBuf<unsigned long long> d_cctr;
        HIP_TRY(hspf.alloc(std::max<uint64_t>(nsurv, 1)));
        HIP_TRY(d_cctr.alloc(5));
        HIP_TRY(hipMemsetAsync(d_cctr.p, 0, sizeof(unsigned long long) * 5, st));
        HIP_TRY(hipMemsetAsync(hspf.p, 0, std::max<uint64_t>(nsurv, 1), st));
        if (nsurv) {
            CoverArgs ca{};
            ca.keys = keys.p;
            ca.n = nsurv;
            ca.kl = kl;
            ca.xres = xres.p;
            ca.hsp_threshold = P->hsp_threshold;
            ca.hsp = hspf.p;
            ca.ctr = d_cctr.p;
            const uint64_t cgrid = std::min<uint64_t>((uint64_t)c->cus * 16, (nsurv + 255) / 256);
            SABER_LAUNCH(coverage_kernel, dim3((uint32_t)std::max<uint64_t>(cgrid, 1)), dim3(256), 0, st, ca);
            HIP_TRY(hipGetLastError());
        }
        lap("coverage");
        // HSP indices in key order
        DevBuf<uint32_t> hidx;
        DevBuf<uint64_t> d_nh;
        HIP_TRY(hidx.alloc(std::max<uint64_t>(nsurv, 1)));
        HIP_TRY(d_nh.alloc(1));
        uint64_t nh = 0;
        if (nsurv) {
            size_t tb = 0;
            rocprim::counting_iterator<uint32_t> idx(0);
            HIP_TRY(rocprim::select(nullptr, tb, idx, hspf.p, hidx.p, d_nh.p, (size_t)nsurv, st));
            HIP_TRY(sort_tmp.reserve(tb));
            HIP_TRY(rocprim::select(sort_tmp.p, tb, idx, hspf.p, hidx.p, d_nh.p, (size_t)nsurv, st));
            HIP_TRY(hipMemcpyAsync(&nh, d_nh.p, sizeof nh, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
        ctr.hsps = nh;
        lap("hsp select");
        DevBuf<saber_hip_region> rslot;
        DevBuf<uint8_t> rkeep;
        HIP_TRY(rslot.alloc(std::max<uint64_t>(nh, 1)));
        HIP_TRY(rsel.alloc(std::max<uint64_t>(nh, 1)));
        HIP_TRY(rkeep.alloc(std::max<uint64_t>(nh, 1)));
        DevBuf<int32_t> d_rthr;
        std::vector<int32_t> rthr(nq);
        for (uint32_t q = 0; q < nq; ++q) rthr[q] = prep[q].region_thr;
        if (int r = upload(d_rthr, rthr.data(), rthr.size(), st)) return r;
        if (nh) {
            RegionArgs ra{};
            ra.keys = keys.p;
            ra.xres = xres.p;
            ra.hidx = hidx.p;
            ra.nh = nh;
            ra.kl = kl;
            ra.join = 2 * P->window_slack;
            ra.hsp_threshold = P->hsp_threshold;
            ra.region_thr = d_rthr.p;
            ra.maxr = P->max_regions_per_subject;
            ra.out = rslot.p;
            ra.keep = rkeep.p;
            ra.ctr = d_cctr.p + 4;
            SABER_LAUNCH(region_kernel, dim3(std::min(grid_for(nh, 256), c->cus * 16)), dim3(256), 0, st, ra);
            HIP_TRY(hipGetLastError());
            size_t tb = 0;
            HIP_TRY(rocprim::select(nullptr, tb, rslot.p, rkeep.p, rsel.p, d_nh.p, (size_t)nh, st));
            HIP_TRY(sort_tmp.reserve(tb));
            HIP_TRY(rocprim::select(sort_tmp.p, tb, rslot.p, rkeep.p, rsel.p, d_nh.p, (size_t)nh, st));
            HIP_TRY(hipMemcpyAsync(&nr, d_nh.p, sizeof nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
        lap("regions");
        unsigned long long cc[5];
        HIP_TRY(hipMemcpyAsync(cc, d_cctr.p, sizeof cc, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        ctr.ungapped_extensions = cc[0];
        ctr.extensions_avoided = cc[1];
        ctr.xdrop_steps = cc[2];
        ctr.scan_groups = cc[3];
        ctr.regions_weak_dropped = cc[4];
    }
    keys.release();
    xres.release();
    sort_tmp.release();
    tm.regions = ms_since(t0);

    // ---- cap, plans, thresholds, prefilter and both scoring passes on the device ----
    std::vector<saber_hip_query_stats> qstats(nq);
    std::vector<int64_t> floor_q(nq), rfloor_q(nq);
    std::vector<int32_t> rthr_q(nq);
    for (uint32_t q = 0; q < nq; ++q) {
        qstats[q] = prep[q].qs;
        floor_q[q] = prep[q].floor;
        rfloor_q[q] = prep[q].region_floor;
        rthr_q[q] = prep[q].region_thr;
    }
    const double ecut = P->evalue > 0.0 && std::isfinite(P->evalue) ? P->evalue : INFINITY;
    SearchScoreIn si{};
    si.st = st;
    si.regions = rsel.p;
    si.nregions = nr;
    si.nq = nq;
    si.d_qoff = d_qoff.p;
    si.d_qlen = d_qlen.p;
    si.d_qres = d_q.p + kGuard;
    si.qtotal = qtotal;
    si.d_qbias = any_bias ? d_bias.p + kGuard : nullptr;
    si.qstats = qstats.data();
    si.floor = floor_q.data();
    si.region_floor = rfloor_q.data();
    si.region_thr = rthr_q.data();
    si.region_e = region_e;
    si.ecut = ecut;
    si.P = P;
    si.scoring = scoring;
    si.prefilter = prefilter;
    si.keep_regions = P->keep_regions != 0;
    SearchScoreOut so;
    if (int r = search_score_device(c, si, &so)) return r;
    rsel.release();
    tm.regions += so.ms_regions;
    tm.plan = so.ms_plan;
    tm.prefilter = so.ms_prefilter;
    tm.score = so.ms_score;
    ctr.regions = so.regions;
    ctr.regions_dropped_by_cap = so.dropped_by_cap;
    ctr.band_scored_regions = so.band_scored;
    ctr.band_rescored = so.rescored;
    ctr.banded_cells = so.banded_cells;
    ctr.regions_pruned = so.pruned;
    const std::vector<uint64_t>& roff = so.roff;
    const uint64_t nreg = so.regions;
    auto region_count = [&](uint32_t q) { return roff[q + 1] - roff[q]; };
    const std::vector<SearchCand>& cand = so.cands;   // in capped order, so grouped by query
    std::vector<uint64_t> cstart(nq + 1, 0);
    for (const SearchCand& x : cand) cstart[x.q + 1]++;
    for (uint32_t q = 0; q < nq; ++q) cstart[q + 1] += cstart[q];
    saber_hip_seqset qs{res->qres.data(), queries->offsets, nq, any_bias ? bias_all.data() : nullptr};

    // ---- finish (engine.rs:1127-1221) ----
    t0 = Clock::now();
    // HitCand::ri indexes the candidates below
    std::vector<std::vector<HitCand>> order(nq);
    parallel_for(nq, [&](uint64_t q) {
        const QPrep& pq = prep[q];
        for (uint64_t i = cstart[q]; i < cstart[q + 1]; ++i) {
            const int64_t s = cand[i].score;
            if (!P->per_subject_evalues) {
                if (s >= pq.floor) order[q].push_back(HitCand{saber_hip_stats_evalue(&pq.qs, s), i});
                continue;
            }
            if (s <= 0) continue;
            const uint64_t n = cand[i].slen;
            if (saber_hip_stats_evalue_in(&pq.qs, s, saber_hip_stats_space_lower_bound(&pq.qs, n)) > ecut) continue;
            const double e = saber_hip_stats_evalue_for(&pq.qs, s, n);
            if (e <= ecut) order[q].push_back(HitCand{e, i});
        }
        std::sort(order[q].begin(), order[q].end(), [&](const HitCand& a, const HitCand& b) {
            if (a.e != b.e) return a.e < b.e;
            if (cand[a.ri].score != cand[b.ri].score) return cand[a.ri].score > cand[b.ri].score;
            return region_before(cand[a.ri].r, cand[b.ri].r);
        });
    });
    for (uint32_t q = 0; q < nq; ++q) ctr.regions_below_evalue += region_count(q) - order[q].size();

    // The reporting loop per query, replayed over tracebacks computed in
    // rounds: each round traces, for every query still open, its next
    // candidates whose subject has not yet reached max_hsps; the loop then
    // consumes them in order. Counters count only the tracebacks the loop
    // asked for.
    struct Traced {
        saber_hip_alignment a;
        uint8_t exact;
        std::vector<uint8_t> oc;
        std::vector<uint32_t> ol;
        saber_hip_traceback_counters tc;
    };
    struct LoopState {
        size_t pos = 0;                       // next candidate in order
        bool done = false;
        uint64_t reported_subjects = 0;
        std::map<uint64_t, uint32_t> per_subject;
        std::map<uint64_t, Traced> traced;    // candidate index in order -> result
    };
    std::vector<LoopState> LS(nq);
    std::vector<uint8_t> reported(cand.size(), 0);
    const uint64_t batch0 = (uint64_t)P->max_target_seqs * P->max_hsps;
    std::vector<uint64_t> want(nq, batch0);
    auto make_hit = [&](uint32_t q, const saber_hip_region& r, uint64_t lo, const Traced& T) {
        const uint8_t* qq = res->qres.data() + queries->offsets[q];
        const uint64_t s0 = c->db_off[r.subject];
        const uint64_t sfull = c->db_off[r.subject + 1] - s0;
        const uint8_t* sp = c->db_host.data() + s0 + lo;
        saber_hip_hit h{};
        uint64_t i = T.a.q_start, j = T.a.s_start;
        uint64_t cols = 0;
        for (uint32_t k = 0; k < T.a.nops; ++k) {
            const uint8_t op = T.oc[k];
            const uint32_t l = T.ol[k];
            cols += l;
            if (op == 'M') {
                for (uint32_t t = 0; t < l; ++t) {
                    const uint8_t a = qq[i], b = sp[j];
                    const bool ident = a == b && a < 20;
                    const bool pos = ident || scoring->matrix[(a < 24 ? a : kCodeX) * 24 + (b < 24 ? b : kCodeX)] > 0;
                    if (ident) h.matches++;
                    else h.mismatches++;
                    h.positives += pos;
                    ++i;
                    ++j;
                }
            } else if (op == 'I') {
                h.gap_opens++;
                h.gap_columns += l;
                i += l;
            } else {
                h.gap_opens++;
                h.gap_columns += l;
                j += l;
            }
        }
        h.subject = r.subject;
        h.score = T.a.score;
        h.bitscore = saber_hip_stats_bitscore(&prep[q].qs, T.a.score);
        h.evalue = saber_hip_stats_evalue_for(&prep[q].qs, T.a.score, sfull);
        h.q_start = T.a.q_start + 1;
        h.q_end = T.a.q_end;
        h.s_start = T.a.s_start + lo + 1;
        h.s_end = T.a.s_end + lo;
        h.columns = cols;
        h.q_len = qlen[q];
        h.s_len = sfull;
        h.exact_in_window = T.exact;
        h.nops = T.a.nops;
        return h;
    };
    for (;;) {
        // candidates to trace this round
        std::vector<saber_hip_traceback_job> jobs;
        std::vector<std::pair<uint32_t, size_t>> job_of;   // (query, index in order)
        for (uint32_t q = 0; q < nq; ++q) {
            LoopState& L = LS[q];
            if (L.done) continue;
            uint64_t taken = 0;
            for (size_t k = L.pos; k < order[q].size() && taken < want[q]; ++k) {
                if (L.traced.count(k)) continue;
                const saber_hip_region& r = cand[order[q][k].ri].r;
                auto it = L.per_subject.find(r.subject);
                if (it != L.per_subject.end() && it->second >= P->max_hsps) continue;
                const uint64_t ri = order[q][k].ri;
                const saber_hip_plan& pl = cand[ri].p;
                saber_hip_traceback_job J{};
                J.query = q;
                J.subject = (uint32_t)r.subject;
                J.win_lo = (uint32_t)pl.lo;
                J.win_hi = (uint32_t)pl.hi;
                J.d_min = r.d_min;
                J.d_max = r.d_max;
                J.best = r.best;
                J.want = cand[ri].score;
                J.has_limit = pl.banded ? 1 : 0;
                J.limit_lo = pl.band_lo;
                J.limit_hi = pl.band_hi;
                jobs.push_back(J);
                job_of.emplace_back(q, k);
                ++taken;
            }
        }
        if (!jobs.empty()) {
            uint64_t opcap = 0;
            for (const auto& J : jobs) opcap += qlen[J.query] + (J.win_hi - J.win_lo);
            std::vector<saber_hip_alignment> al(jobs.size());
            std::vector<uint8_t> ex(jobs.size());
            std::vector<uint8_t> oc(std::max<uint64_t>(opcap, 1));
            std::vector<uint32_t> ol(std::max<uint64_t>(opcap, 1));
            std::vector<saber_hip_traceback_counters> jc(jobs.size());
            c->mu.unlock();
            int r = saber_hip_traceback_ex(c, &qs, nullptr, jobs.data(), jobs.size(), scoring, P->band_pad, 0,
                                           P->traceback_max_bytes, al.data(), ex.data(), oc.data(), ol.data(), opcap,
                                           nullptr, jc.data());
            c->mu.lock();
            if (r) return r;
            for (size_t x = 0; x < jobs.size(); ++x) {
                Traced T;
                T.a = al[x];
                T.exact = ex[x];
                T.tc = jc[x];
                if (al[x].status == 0 && al[x].nops) {
                    T.oc.assign(oc.begin() + al[x].op_off, oc.begin() + al[x].op_off + al[x].nops);
                    T.ol.assign(ol.begin() + al[x].op_off, ol.begin() + al[x].op_off + al[x].nops);
                }
                LS[job_of[x].first].traced.emplace(job_of[x].second, std::move(T));
            }
        }
        // replay the loops: each query's loop, hits and counters are its own
        std::vector<uint8_t> qopen(nq, 0);
        std::vector<saber_hip_search_counters> qc(nq);
        parallel_for(nq, [&](uint64_t qq) {
            const uint32_t q = (uint32_t)qq;
            LoopState& L = LS[q];
            if (L.done) return;
            saber_hip_search_counters& ctr = qc[q];
            auto& hits = res->hits[q];
            while (L.pos < order[q].size()) {
                const uint64_t ri = order[q][L.pos].ri;
                const saber_hip_region& r = cand[ri].r;
                const uint64_t subj = r.subject;   // subject_group = 1
                auto it = L.per_subject.find(subj);
                const uint32_t cnt = it == L.per_subject.end() ? 0 : it->second;
                if (L.reported_subjects >= P->max_target_seqs && it == L.per_subject.end()) {
                    L.done = true;
                    break;
                }
                if (cnt >= P->max_hsps) {
                    ++L.pos;
                    continue;
                }
                auto tt = L.traced.find(L.pos);
                if (tt == L.traced.end()) break;   // needs the next round
                const Traced& T = tt->second;
                ctr.tracebacks += T.tc.tracebacks;
                ctr.band_widenings += T.tc.band_widenings;
                ctr.band_capped += T.tc.band_capped;
                ctr.end_anchored_tracebacks += T.tc.end_anchored_tracebacks;
                ctr.banded_cells += T.tc.banded_cells;
                ++L.pos;
                if (T.a.status != 0) continue;   // no alignment
                saber_hip_hit h = make_hit(q, r, cand[ri].p.lo, T);
                bool overlap = false;
                if (cnt > 0)
                    for (const auto& o : hits)
                        if (o.subject == h.subject && o.q_start <= h.q_end && h.q_start <= o.q_end &&
                            o.s_start <= h.s_end && h.s_start <= o.s_end)
                            overlap = true;
                if (overlap) continue;
                const bool passes = P->per_subject_evalues ? h.evalue <= ecut : (int64_t)h.score >= prep[q].floor;
                const double ident = h.columns ? 100.0 * (double)h.matches / (double)h.columns : 0.0;
                const double qcov = h.q_len ? 100.0 * (double)(h.q_end + 1 - h.q_start) / (double)h.q_len : 0.0;
                const double scov = h.s_len ? 100.0 * (double)(h.s_end + 1 - h.s_start) / (double)h.s_len : 0.0;
                if (ident < P->min_identity || qcov < P->min_query_cover || scov < P->min_subject_cover || !passes) {
                    ctr.filtered_identity_coverage++;
                    continue;
                }
                if (cnt == 0) L.reported_subjects++;
                L.per_subject[subj] = cnt + 1;
                h.op_off = res->op_codes[q].size();
                res->op_codes[q].insert(res->op_codes[q].end(), T.oc.begin(), T.oc.end());
                res->op_lens[q].insert(res->op_lens[q].end(), T.ol.begin(), T.ol.end());
                hits.push_back(h);
                reported[ri] = 1;
            }
            if (L.pos >= order[q].size()) L.done = true;
            if (!L.done) {
                qopen[q] = 1;
                want[q] *= 2;
            }
        });
        bool open = false;
        for (uint32_t q = 0; q < nq; ++q) {
            const saber_hip_search_counters& x = qc[q];
            ctr.tracebacks += x.tracebacks;
            ctr.band_widenings += x.band_widenings;
            ctr.band_capped += x.band_capped;
            ctr.end_anchored_tracebacks += x.end_anchored_tracebacks;
            ctr.banded_cells += x.banded_cells;
            ctr.filtered_identity_coverage += x.filtered_identity_coverage;
            open = open || qopen[q];
        }
        if (!open) break;
    }
    parallel_for(nq, [&](uint64_t qq) {
        const uint32_t q = (uint32_t)qq;
        auto& hits = res->hits[q];
        // the operations move with their hits
        std::vector<size_t> idx(hits.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            const saber_hip_hit &x = hits[a], &y = hits[b];
            if (x.evalue != y.evalue) return x.evalue < y.evalue;
            if (x.score != y.score) return x.score > y.score;
            if (x.subject != y.subject) return x.subject < y.subject;
            return x.s_start < y.s_start;
        });
        std::vector<saber_hip_hit> sorted;
        std::vector<uint8_t> oc;
        std::vector<uint32_t> ol;
        for (size_t k : idx) {
            saber_hip_hit h = hits[k];
            const uint64_t o = h.op_off;
            h.op_off = oc.size();
            oc.insert(oc.end(), res->op_codes[q].begin() + o, res->op_codes[q].begin() + o + h.nops);
            ol.insert(ol.end(), res->op_lens[q].begin() + o, res->op_lens[q].begin() + o + h.nops);
            sorted.push_back(h);
        }
        hits.swap(sorted);
        res->op_codes[q].swap(oc);
        res->op_lens[q].swap(ol);
    });
    for (uint32_t q = 0; q < nq; ++q) ctr.alignments += res->hits[q].size();
    tm.finish = ms_since(t0);

    // ---- the regions as the caller may inspect them (keep_regions) ----
    if (P->keep_regions) {
        std::vector<double> reg_e(nreg, INFINITY);
        std::vector<uint8_t> rep(nreg, 0);
        for (uint32_t q = 0; q < nq; ++q)
            for (const HitCand& h : order[q]) {
                reg_e[cand[h.ri].ri] = h.e;
                rep[cand[h.ri].ri] = reported[h.ri];
            }
        for (uint32_t q = 0; q < nq; ++q) {
            auto& R = res->regions[q];
            R.resize(region_count(q));
            for (uint64_t i = roff[q]; i < roff[q + 1]; ++i) {
                saber_hip_search_region& o = R[i - roff[q]];
                const saber_hip_region& r = so.flat[i];
                const saber_hip_plan& pl = so.plans[i];
                o.subject = r.subject;
                o.d_min = r.d_min;
                o.d_max = r.d_max;
                o.best = r.best;
                o.n_hsp = r.n_hsp;
                o.hsp_sum = r.hsp_sum;
                o.score = so.scores[i];
                o.win_lo = pl.lo;
                o.win_hi = pl.hi;
                o.band_lo = pl.band_lo;
                o.band_hi = pl.band_hi;
                o.banded = pl.banded;
                o.flags = (so.rescored_flag[i] ? 1u : 0u) | (std::isfinite(reg_e[i]) ? 2u : 0u) | (rep[i] ? 4u : 0u) |
                          (so.pruned_flag[i] ? 8u : 0u);
                o.evalue = reg_e[i];
            }
        }
        res->flat.assign(so.flat.begin(), so.flat.end());
        res->plans.assign(so.plans.begin(), so.plans.end());
        res->qidx.assign(so.qidx.begin(), so.qidx.end());
        res->slen.assign(so.slen.begin(), so.slen.end());
    }
    res->kept_regions = P->keep_regions != 0;
    res->qlen = qlen;
    res->rthr = rthr_q;
    res->db_gen = c->db_gen;
    tm.total = ms_since(t_start);
    *out = res.release();
    return SABER_HIP_OK;
}

uint64_t saber_hip_search_hit_count(const saber_hip_search* s, uint32_t q) {
    return s && q < s->nq ? s->hits[q].size() : 0;
}

int saber_hip_search_hits(const saber_hip_search* s, uint32_t q, saber_hip_hit* out) {
    if (!s || q >= s->nq) return fail(SABER_HIP_EINVAL, "bad search or query");
    if (!s->hits[q].empty() && !out) return fail(SABER_HIP_EINVAL, "out is null");
    std::copy(s->hits[q].begin(), s->hits[q].end(), out);
    return SABER_HIP_OK;
}

uint64_t saber_hip_search_op_count(const saber_hip_search* s, uint32_t q) {
    return s && q < s->nq ? s->op_codes[q].size() : 0;
}

int saber_hip_search_ops(const saber_hip_search* s, uint32_t q, uint8_t* codes, uint32_t* lens) {
    if (!s || q >= s->nq) return fail(SABER_HIP_EINVAL, "bad search or query");
    if (!s->op_codes[q].empty() && (!codes || !lens)) return fail(SABER_HIP_EINVAL, "null output");
    std::copy(s->op_codes[q].begin(), s->op_codes[q].end(), codes);
    std::copy(s->op_lens[q].begin(), s->op_lens[q].end(), lens);
    return SABER_HIP_OK;
}

int saber_hip_search_all_hits(const saber_hip_search* s, uint64_t* hit_off, saber_hip_hit* out, uint64_t* op_base,
                              uint8_t* codes, uint32_t* lens) {
    if (!s) return fail(SABER_HIP_EINVAL, "search is null");
    uint64_t h = 0, o = 0;
    for (uint32_t q = 0; q < s->nq; ++q) {
        if (hit_off) hit_off[q] = h;
        if (op_base) op_base[q] = o;
        if (out) std::copy(s->hits[q].begin(), s->hits[q].end(), out + h);
        if (codes) std::copy(s->op_codes[q].begin(), s->op_codes[q].end(), codes + o);
        if (lens) std::copy(s->op_lens[q].begin(), s->op_lens[q].end(), lens + o);
        h += s->hits[q].size();
        o += s->op_codes[q].size();
    }
    if (hit_off) hit_off[s->nq] = h;
    if (op_base) op_base[s->nq] = o;
    return SABER_HIP_OK;
}

uint64_t saber_hip_search_region_count(const saber_hip_search* s, uint32_t q) {
    return s && q < s->nq ? s->regions[q].size() : 0;
}

int saber_hip_search_regions(const saber_hip_search* s, uint32_t q, saber_hip_search_region* out) {
    if (!s || q >= s->nq) return fail(SABER_HIP_EINVAL, "bad search or query");
    if (!s->kept_regions) return fail(SABER_HIP_ESTATE, "the search kept no regions (keep_regions = 0)");
    if (!s->regions[q].empty() && !out) return fail(SABER_HIP_EINVAL, "out is null");
    std::copy(s->regions[q].begin(), s->regions[q].end(), out);
    return SABER_HIP_OK;
}

int saber_hip_search_features(saber_hip_search* s, const float* mean, const float* inv_std, uint16_t* out) {
    if (!s || !s->ctx) return fail(SABER_HIP_EINVAL, "null search");
    if (!s->kept_regions) return fail(SABER_HIP_ESTATE, "the search kept no regions (keep_regions = 0)");
    const uint64_t n = s->flat.size();
    if (n == 0) return SABER_HIP_OK;
    if (!out) return fail(SABER_HIP_EINVAL, "out is null");
    saber_hip_ctx* c = s->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    if (c->db_gen != s->db_gen) return fail(SABER_HIP_ESTATE, "the database changed since the search");
    hipStream_t st = c->stream;
    DevBuf<uint8_t> d_q;
    DevBuf<uint64_t> d_qoff;
    const uint64_t qtotal = s->qoff.back();
    if (int r = upload_codes(d_q, s->qres.data(), qtotal, st)) return r;
    if (int r = upload(d_qoff, s->qoff.data(), s->qoff.size(), st)) return r;
    DevBuf<__bf16> X;
    if (int r = region_features_dev(c, st, d_qoff.p, d_q.p + kGuard, s->nq, s->flat.data(), s->flat.size(),
                                    s->plans.data(), s->qidx.data(), s->qlen.data(), s->slen.data(), s->rthr.data(),
                                    mean, inv_std, X))
        return r;
    HIP_TRY(hipMemcpyAsync(out, X.p, 2 * 64 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}

int saber_hip_search_get_counters(const saber_hip_search* s, saber_hip_search_counters* out) {
    if (!s || !out) return fail(SABER_HIP_EINVAL, "null argument");
    *out = s->ctr;
    return SABER_HIP_OK;
}

int saber_hip_search_get_timing(const saber_hip_search* s, saber_hip_search_timing* out) {
    if (!s || !out) return fail(SABER_HIP_EINVAL, "null argument");
    *out = s->tm;
    return SABER_HIP_OK;
}

void saber_hip_search_destroy(saber_hip_search* s) { delete s; }

}  // extern "C"
