// Shared device/host definitions for the SABER MI355X scoring library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/saber_hip.h"
#include "align_limits.h"

namespace saber {

// Alignment codes 0..23 (ARNDCQEGHILKMFPSTWYVBZX*); any code >= 24 scores as
// X (src/alphabet.rs:39-41), so residues are stored canonical on upload.
constexpr int kCodeX = 22;

// An unaligned dword load from global memory issued as a global (not flat)
// access. A generic pointer that reaches a non-inlined function is loaded
// with flat instructions, which count against the LDS wait counter as well:
// every wait for an LDS read would then also wait for this load.
__device__ __forceinline__ uint32_t gload_u32(const void* p) {
    struct __attribute__((packed)) U32 { uint32_t v; };
    return ((const __attribute__((address_space(1))) U32*)p)->v;
}

// A byte pointer into LDS that keeps its address space across non-inlined
// calls (a generic one is read there with flat loads), and an aligned dword
// read through it.
typedef const __attribute__((address_space(3))) char* lds_bytes;
__device__ __forceinline__ uint32_t lds_u32(lds_bytes p) {
    return *(const __attribute__((address_space(3))) uint32_t*)p;
}
// A kernel launch checked by the hipGetLastError that follows it: the error
// state is cleared first, so that a failed HIP call tolerated here (or made by
// other code on this thread) is not reported as the launch's failure.
#define SABER_LAUNCH(...)                  \
    do {                                   \
        (void)hipGetLastError();           \
        hipLaunchKernelGGL(__VA_ARGS__);   \
    } while (0)

constexpr short kMax16 = 32767;      // any job whose best reaches this is rescored in int32

// Band kernel score table: int8, 25 query rows x 25 columns (subject codes
// 0..23 and the sentinel column 24, scoring -128; query row 24 is the dead
// row, scoring -128 against every column), staged into LDS as words. Code 24
// stands for a subject column outside the window and for a query row outside
// the query.
constexpr int kBandSentinel = 24;
constexpr uint32_t kSentinel4 = 0x18181818u;  // four sentinel bytes
constexpr uint32_t kTable8Stride = 25;
constexpr int kTable8Words = (25 * 25 + 3) / 4;
constexpr uint32_t kMaxBandRows = 32767;      // rows of one packed job (row index kept in 16 bits)

// One banded job after planning (band clipped to its rectangle, rows that
// hold band cells). Band cell k of row i is subject column j = i + lo + k.
// A swapped job is the transposed problem of a scores-only job without bias
// under a symmetric matrix (same score): its rows run over the subject window
// (q_off is in the residue blob) and its columns over the query (s_off is in
// the query blob), band negated.
struct BandJobDev {
    uint64_t q_off;     // first query residue (and bias entry) in the query blob
    uint64_t s_off;     // first subject-window residue in the residue blob
    uint32_t m : 31;    // query length
    uint32_t swapped : 1;
    uint32_t n;         // window length
    int32_t lo;         // clipped band lo (diagonal j - i)
    int32_t w;          // clipped band width
    int32_t i_first;    // first row holding a band cell
    int32_t rows;       // number of rows holding band cells
};

struct BandOutputs {
    int32_t* score;
    uint32_t* q_end;
    uint32_t* s_end;
    uint64_t* cells;
};

// One work chunk of the multi-shape band launch (one wave): pairs
// [pair_off, pair_off + npairs) of the pair list, scored with shape `shape`;
// flags: kChunkShared when both jobs of every pair share their query rows.
struct BandBlock {
    uint32_t shape, pair_off, npairs, flags;
};
constexpr uint32_t kChunkShared = 1u;

// Pair table of the shared-query flavour: for each query code q <= 24 (24 =
// the dead row) and subject codes a, b <= 24 (24 = sentinel), the dword
// (S(q, a), S(q, b)) as two int16, at byte (q * 625 + a * 25 + b) * 4.
constexpr uint32_t kPairRowBytes = 625 * 4;
constexpr int kPairWords = 25 * 625;

struct BandKernelArgs {
    const uint32_t* pairs;      // 2 job ids per pair (0xFFFFFFFF = none)
    uint32_t npairs;
    const BandBlock* blocks;    // work chunks in dispatch order
    uint32_t nchunks;
    uint32_t* next_chunk;       // claim counter (zero before the launch)
    const uint32_t* pair_table; // kPairWords words (global; staged into LDS)
    const BandJobDev* jobs;
    const uint8_t* qres;        // device buffers with 16 guard bytes before and
    const int8_t* qbias;        // 64 after (qbias may be null without bias)
    const uint8_t* sres;
    uint64_t qres_len, sres_len;
    const uint32_t* table;      // int8 band table as words (global; staged into LDS)
    int32_t go, ge;
    BandOutputs out;            // score, q_end, s_end (cells come from the planner)
    uint32_t* flagged;          // job ids needing the exact int32 rescore
    uint32_t* nflagged;
};

// Generic exact int32 kernel over a list of job ids.
struct GenericArgs {
    const uint32_t* ids;
    uint32_t nids;
    const BandJobDev* jobs;
    const uint8_t* qres;
    const int8_t* qbias;
    const uint8_t* sres;
    const int32_t* mat;         // 24 x 24 int32
    int32_t go, ge;
    BandOutputs out;
    int32_t* work;              // per id: 2 * (w + 1) int32 at work_off[id]
    const uint64_t* work_off;
};

// One compiled variant of the packed band kernel: L lanes per job pair,
// C band cells per lane, K capped trailing cells per lane (band width W = L*C
// covers widths W-K+1 .. W, every width up to W when K = C), with or without
// the query bias. Variant v is shape v / 2 of the multi-shape kernel.
struct BandVariant {
    int L, C, K;
    int bias;
};

const BandVariant* band_variants(int* count);
hipError_t launch_band_multi(const BandKernelArgs& a, uint32_t nworkgroups, bool bias, bool ends, hipStream_t s);
hipError_t launch_generic(const GenericArgs& a, hipStream_t s);
// a.ids[0 .. *count) with the count read on the device; nthreads threads (a
// multiple of 128), each with slot_ints int32 of a.work (a.work_off unused).
hipError_t launch_generic_slots(const GenericArgs& a, const uint32_t* count, uint32_t nthreads, uint32_t slot_ints,
                                hipStream_t s);
// Full-matrix (query, subject) pairs scored exactly, count read on the device.
hipError_t launch_full_int32(const GenericArgs& a, const uint32_t* pairs, const uint32_t* count, const uint64_t* qoff,
                             const uint64_t* soff, uint32_t nsub, int32_t* score, uint32_t* q_end, uint32_t* s_end,
                             uint32_t nthreads, uint32_t slot_ints, hipStream_t s);

}  // namespace saber

namespace saber {

// X-drop score table: 32 query rows x 256 subject columns, indexed by
// (q << 8 | s); rows and columns past the 24 codes hold X scores.
constexpr int kXdropTable = 32 * 256;
// hit ranges / claim counters of the X-drop kernel (counters 256 B apart)
constexpr uint32_t kXdropGroups = 8;
constexpr uint32_t kXdropCounterStride = 64;

struct XdropArgs {
    const uint32_t* hits;      // 4 u32 per hit: query, subject, qp, sp
    uint32_t nhits;
    uint32_t chunk;            // hits per claim
    const uint64_t* qoff;      // query offsets (count+1)
    const uint8_t* qres;
    const int8_t* qbias;       // may be null
    const uint64_t* soff;      // subject offsets
    const uint8_t* sres;
    const short* table;        // kXdropTable entries
    int32_t x;
    uint32_t* next_chunk;      // kXdropGroups claim counters, kXdropCounterStride apart (cleared before launch)
    void* desc;                // 32 B per hit, written by the descriptor pass
    void* res;                 // 16 B per hit: score, q_lo, q_hi, steps
};

// Workgroups of the X-drop kernel resident per CU (its grid is CUs x this).
int xdrop_blocks_per_cu(bool bias, bool wide);
// wide: some score lies outside int8 (int16 table in LDS).
hipError_t launch_xdrop(const XdropArgs& a, bool bias, bool wide, int grid, hipStream_t s);
// The X-drop kernel alone, over descriptors already in a.desc (32 B per hit:
// absolute query / subject positions of the seed, its query position, the
// steps available to the right including the seed and to the left).
hipError_t launch_xdrop_described(const XdropArgs& a, bool bias, bool wide, int grid, hipStream_t s);
struct alignas(16) XdropDescDev {
    uint64_t q_abs, s_abs;
    uint32_t qp, n_right, n_left, pad;
};

}  // namespace saber

namespace saber {

// Full-matrix (exhaustive) kernel over the cross product of queries and
// subjects; outputs are [query slot][subject].
// Banded traceback (align.hip): one record per job.
struct AlignRecord {
    int32_t status;             // 0 alignment, 1 none (empty band or best <= 0)
    int32_t score;
    uint32_t q_start, q_end, s_start, s_end;
    uint32_t nops;              // run-length operations at the job's op offset
    uint32_t pad;
    uint64_t cells;
};

struct AlignArgs {
    const uint32_t* ids;        // jobs of this launch
    uint32_t n;
    const BandJobDev* jobs;
    const uint8_t* qres;
    const int8_t* qbias;
    const uint8_t* sres;
    const int32_t* mat;         // 24 x 24 int32
    int32_t go, ge;
    const int32_t* stop_at;     // per job, or null
    const uint64_t* work_off;   // per launch slot: 2 * (w + 1) int32 (one-thread jobs)
    int32_t* work;
    const uint64_t* dir_off;    // per launch thread: rows * w bytes
    uint8_t* dir;
    const uint64_t* op_off;     // per launch thread: m + n entries
    uint8_t* op_codes;
    uint32_t* op_lens;
    AlignRecord* out;           // indexed by job id
};

// One thread per job, rows in the work buffer (band widths over 16 * 64).
hipError_t launch_band_align(const AlignArgs& a, hipStream_t s);
hipError_t launch_gather_ops(uint32_t n, const uint32_t* ids, const AlignRecord* recs, const uint64_t* op_off,
                             const uint64_t* dst_off, const uint8_t* codes, const uint32_t* lens, uint8_t* dcodes,
                             uint32_t* dlens, hipStream_t s);
// Lane classes of one launch: the launch's slots and blocks of each class, as
// prefix offsets, and the block order: launch block b runs block
// order[b] & 0xFFFFFF of class order[b] >> 24.
struct AlignLaneClasses {
    uint32_t slot_start[kAlignLaneClasses + 1];
    uint32_t block_start[kAlignLaneClasses + 1];
    const uint32_t* order;
};
hipError_t launch_band_align_lanes(const AlignArgs& a, const AlignLaneClasses& cls, hipStream_t s);

constexpr int kMaxFullSegs = 32;   // segments per register class (shapes)

// A segment of the full-matrix chunk list: one shape's subject pairs against
// every query; chunk ch of query rank qr is chunk start + qr * nch + ch.
struct FullSeg {
    uint32_t start;             // first chunk
    uint32_t shape;             // shape id in the flavour's list
    uint32_t nch;               // chunks per query (64 / L pairs each)
    uint32_t pair_off;          // first pair of the shape in the pair list
    uint32_t npairs;
    uint32_t pad[3];
};

struct FullKernelArgs {
    const uint32_t* pairs;      // 2 subject ids per pair (0xFFFFFFFF = none)
    const FullSeg* segs;        // this register class's segments
    uint32_t nsegs;
    uint32_t nchunks;
    uint32_t* next_chunk;       // claim counter (zero before the launch)
    const uint32_t* qorder;     // query ids by rank (longest first)
    const uint64_t* qoff;       // query offsets
    const uint8_t* qres;        // guard-padded like the band kernel's buffers
    const int8_t* qbias;
    uint64_t qres_len;
    const uint64_t* soff;       // subject offsets
    const uint8_t* sres;
    const uint32_t* pair_table; // kPairWords words (global; staged into LDS)
    int32_t go, ge;
    uint32_t nsub;              // row stride of the outputs (subjects per query)
    int32_t* score;
    uint32_t* q_end;
    uint32_t* s_end;
    uint32_t* flagged;          // (query, subject) pairs needing the int32 rescore
    uint32_t* nflagged;
};

// A compiled full-matrix shape: L lanes per subject pair, C columns per lane;
// nt = threads per block of its register class.
struct FullShape {
    int L, C, nt;
};

// The shapes of a flavour (ends: with end cells), in shape-id order.
const FullShape* full_shapes(bool ends, int* count);
// One register class's persistent kernel (nt threads per block).
hipError_t launch_full_multi(bool bias, bool ends, int nt, const FullKernelArgs& a, uint32_t grid, hipStream_t s);
// Blocks of that kernel resident per CU.
int full_multi_blocks_per_cu(bool bias, bool ends, int nt);

}  // namespace saber

namespace saber {

// Prefilter MLP (mlp.hip). Weights in the operand layouts described there.
struct MlpWeights {
    const __bf16* W1T;
    const __bf16* W2P;
    const __bf16* W2Q;
    const float* b1;
    const float* b2;
    const float* w3;
    const float* b3;   // [1]
};

struct MlpGrads {
    float* dW1;    // [64][256]
    float* dW2;    // [256][256]
    float* db1;    // [256]
    float* db2;    // [256]
    float* dw3;    // [256]
    float* db3;    // [1]
    float* loss;   // [1]
};

hipError_t mlp_launch_forward(const MlpWeights& w, const __bf16* X, uint64_t n, float* logits,
                              int cus, hipStream_t s);
hipError_t mlp_launch_backward_act(const MlpWeights& w, const __bf16* X, const float* labels,
                                   uint64_t n, float pos_weight, float inv_n, const MlpGrads& gr,
                                   __bf16* H1, __bf16* H2, __bf16* dH2, __bf16* dH1, float* dy,
                                   int grid, hipStream_t s);
hipError_t mlp_launch_colsum(const __bf16* M, const float* wgt, uint64_t n, float* out, hipStream_t s);
hipError_t mlp_launch_gemm_tn(const __bf16* A, int lda, int a_dim, const __bf16* B, int ldb, int b_dim,
                              float* C, uint64_t K, int kchunks, hipStream_t s);
hipError_t mlp_launch_adam(float* p, float* grad, float* m, float* v, uint32_t n, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t mlp_launch_pack(const float* W1, const float* W2, __bf16* W1T, __bf16* W2P, __bf16* W2Q,
                           hipStream_t s);
hipError_t mlp_launch_to_bf16(const float* src, __bf16* dst, uint64_t n, hipStream_t s);

// Prefilter region features (features.hip): region r = jobs[r], hits
// r * per_region .. of hits (4 u32: query, subject, qp, sp) with their X-drop
// results; out [n][64] bf16.
struct FeatureArgs {
    const saber_hip_band_job* jobs;
    uint64_t n;
    const uint64_t* qoff;
    const uint8_t* qres;
    const uint64_t* soff;
    const uint8_t* sres;
    const uint32_t* hits;     
    const int32_t* xscore;
    const uint32_t* xlo;
    const uint32_t* xhi;
    uint32_t per_region;
    int32_t hsp_threshold;
    const float* mean;         // optional standardisation: (f - mean) * inv_std
    const float* inv_std;
    __bf16* out;
    float* qcomp;              // work: composition of each query (20 per query)
};
// Query compositions (nq queries), then one lane per region.
hipError_t launch_region_features(const FeatureArgs& a, uint32_t nq, hipStream_t s);

// A search region as the search's feature kernel reads it.
struct SearchFeatRegion {
    uint64_t s_off;          // window start in the database residues
    uint32_t m, n;           // query and window lengths
    int32_t band_lo, band_hi;// band in window diagonals (the full rectangle when unbanded)
    int32_t best, hsp_sum;
    uint32_t n_hsp, query;
    int32_t d_span;          // d_max - d_min
    float thr;               // the query's region threshold
    uint32_t slen;           // subject length
    uint32_t banded;
};
struct SearchFeatureArgs {
    const SearchFeatRegion* regions;
    uint64_t n;
    const uint8_t* sres;
    const float* mean;
    const float* inv_std;
    __bf16* out;
    float* qcomp;            // work: 20 per query
};
hipError_t launch_search_features(const SearchFeatureArgs& a, const uint64_t* qoff, const uint8_t* qres, uint32_t nq,
                                  hipStream_t s);

}  // namespace saber
