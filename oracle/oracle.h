/*
 * CPU oracle for the candidate-scoring hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's scalar algorithms (SABER 1.4.0,
 * raw-lab/saber). Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load this library, and only as the checker; the product path
 * (saber_amd/, libsaber_hip.so) never links or calls it.
 *
 * Conventions shared with the reference (src/gpu/mod.rs:9-31):
 *   residues are alignment codes 0..23 (ARNDCQEGHILKMFPSTWYVBZX*), any code
 *   >= 24 scores as X (22); a gap of length k costs open + (k-1)*extend; an
 *   optional bias[i] is added to every aligned column of query position i.
 *   Matrices are 24x24 int32, row-major (576 entries).
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   or_sw_reference       brute-force enumeration of all local alignments on
 *                         tiny inputs; reference known answers (WAGNER=37 at
 *                         (6,6), Wx4000=44000, Wx3500=38500, nucleotide 20/8);
 *                         the reference's own CUDA kernel source compiled as
 *                         C++ (oracle/_ref, when /root/reference is present).
 *   or_banded_score       full band == or_sw_reference incl. end cell; band
 *                         score <= full score; explicit (i,j)-coordinate DP
 *                         with -inf outside the band (independent formulation);
 *                         Wx300 band +-5 = 3300 known answer.
 *   or_xdrop              closed form with x -> infinity (best prefix / suffix
 *                         sums), brute-force bound best_ungapped_on_diagonal,
 *                         and a direct restatement of the scalar twin
 *                         (engine.rs:2460-2494) that must agree.
 *   or_band_align         its score, end cell and cells equal or_banded_score
 *                         (full band: or_sw_reference); the CIGAR rescores
 *                         (an independent column-by-column evaluation in the
 *                         tests, reference.rs:58-105) to the score and ends
 *                         at the end cell; every column stays inside the band;
 *                         stopping at the known score gives the same
 *                         alignment with no more cells (banded.rs:734-758);
 *                         the memory cap is reported (banded.rs:760-767);
 *                         Wx4000 in band +-20 = 44 000 as 4000M.
 *   or_composition_bias   the reference's invariants (non-standard residues
 *                         get 0; poly-Q <= -3; diverse mean < 1.5) plus an
 *                         O(n*w) direct evaluation of the formula.
 */
#ifndef SABER_ORACLE_H
#define SABER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/alphabet.rs:39-41 */
int or_code_index(uint8_t c);

/* src/algorithm/reference.rs:15-52: full Gotoh SW in int64, first row-major
 * optimum. Writes score and exclusive ends. */
void or_sw_reference(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                     const int32_t* mat, const int8_t* bias, int64_t go, int64_t ge,
                     int64_t* score, int64_t* q_end, int64_t* s_end);

/* src/algorithm/banded.rs:123-224 (scalar kernel) with the dispatcher's result
 * convention of banded.rs:408-449 (score <= 0 -> zero ends, cells kept). */
void or_banded_score(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                     const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge,
                     int64_t band_lo, int64_t band_hi,
                     int32_t* score, int64_t* q_end, int64_t* s_end, uint64_t* cells);

/* Banded SW with traceback: src/algorithm/banded.rs:227-333
 * (banded_align_scalar: direction bits org | e_ext << 2 | f_ext << 3, ties
 * diagonal > E > F, extension only when strictly better) with the row stop of
 * the AVX2 kernels the dispatcher runs (banded_simd.rs:324, 594: rows after
 * the first one whose running best reaches stop_at are not computed), then
 * walk (banded.rs:336-403). Returns 0 with an alignment, 1 when there is none
 * (empty input, empty band, best score <= 0), 2 when m * w direction bytes
 * exceed max_dir_bytes. ops: run-length codes 'M' / 'I' / 'D' in alignment
 * order; capacity m + n runs. */
int or_band_align(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n, const int8_t* bias,
                  const int32_t* mat, int32_t go, int32_t ge, int64_t band_lo, int64_t band_hi,
                  int32_t stop_at, uint64_t max_dir_bytes, int32_t* score, int64_t* q_start,
                  int64_t* q_end, int64_t* s_start, int64_t* s_end, uint64_t* cells,
                  uint8_t* op_codes, uint32_t* op_lens, int64_t* nops);

/* Independent check of or_banded_score: the same local DP written in (i, j)
 * matrix coordinates with -inf outside lo <= j-i <= hi (O(m*n) memory). */
void or_banded_score_ij(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                        const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge,
                        int64_t band_lo, int64_t band_hi,
                        int32_t* score, int64_t* q_end, int64_t* s_end);

/* Batched form used by tests and the CPU baseline: job k scores
 * q[qoff[k] .. qoff[k]+qlen[k]) against s[soff[k] .. soff[k]+slen[k]) in band
 * [lo[k], hi[k]] with bias (NULL or indexed like q). */
void or_banded_many(const uint8_t* qres, const uint8_t* sres, const int8_t* bias,
                    const uint64_t* qoff, const uint32_t* qlen,
                    const uint64_t* soff, const uint32_t* slen,
                    const int32_t* lo, const int32_t* hi, int64_t njobs,
                    const int32_t* mat, int32_t go, int32_t ge,
                    int32_t* score, uint32_t* q_end, uint32_t* s_end, uint64_t* cells,
                    int nthreads);

/* src/search/engine/scan.rs:859-904 (xdrop_cat; identical results to
 * ungapped_xdrop_prof engine.rs:2425-2455). Right extension includes the seed
 * pair (qp, sp); left starts at (qp-1, sp-1). Returns score, half-open query
 * range [q_lo, q_hi) and the step count of xdrop_cat. */
void or_xdrop(const uint8_t* q, int64_t m, const int8_t* bias, const uint8_t* s, int64_t n,
              const int32_t* mat, int64_t qp, int64_t sp, int32_t x,
              int32_t* score, int64_t* q_lo, int64_t* q_hi, uint64_t* steps);

/* Direct restatement of the scalar twin ungapped_xdrop (engine.rs:2460-2494),
 * used only to cross-check or_xdrop. */
void or_xdrop_scalar_twin(const uint8_t* q, int64_t m, const int8_t* bias, const uint8_t* s,
                          int64_t n, const int32_t* mat, int64_t qp, int64_t sp, int32_t x,
                          int32_t* score, int64_t* q_lo, int64_t* q_hi);

/* src/algorithm/reference.rs:106-119 */
int64_t or_best_ungapped_on_diagonal(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                                     int64_t d, const int32_t* mat, const int8_t* bias);

/* src/scoring/bias.rs:56-122 */
void or_composition_bias(const uint8_t* q, int64_t n, const int32_t* mat, int8_t* out);


/* ---- Scan schedule (oracle/scan.c): query batch -> candidate regions ---- */
typedef struct or_scan_params {
    const uint8_t* alpha_map;   /* [24]: reduced letter per alignment code, 255 = not seedable */
    uint32_t asize;             /* letters of the seed alphabet */
    uint32_t nshapes;
    const uint16_t* offsets;    /* [nshapes][16]: care positions of each shape */
    const uint32_t* weight;     /* [nshapes] */
    const uint32_t* span;       /* [nshapes] */
    uint32_t min_ids, kx;       /* window filter (0 = off), exact-identity weight */
    int32_t x, hsp_threshold;   /* X-drop, HSP threshold */
    int64_t window_slack;
    uint32_t max_regions_per_subject;
    uint32_t pad;
    uint64_t cap;               /* max_candidates */
    uint64_t seed_max_occ;      /* 0 = auto_occ_cap */
    uint64_t ref_db_residues;
} or_scan_params;

typedef struct or_scan_counters {
    uint64_t seed_positions, seeds_frequent_skipped, scan_candidates, hits, hits_filtered;
    uint64_t ungapped_extensions, extensions_avoided, xdrop_steps, hsps, regions_weak_dropped;
    uint64_t regions_dropped_by_cap, regions, scan_groups;
} or_scan_counters;

/* A candidate region with its HSP evidence (count, saturating sum). */
typedef struct or_region {
    uint64_t subject;
    int64_t d_min, d_max;
    int32_t best;
    uint32_t n_hsp;
    int32_t hsp_sum;
    uint32_t pad;
} or_region;

/* auto_occ_cap, src/search/index.rs:329-336 */
uint64_t or_auto_occ_cap(uint64_t total, uint64_t ncodes, uint64_t db_residues, uint64_t ref_residues);
/* Engine::search_scan up to the per-query capped region lists. qbias: NULL
 * or one int8 per query residue; region_thr[q]: the query's region
 * threshold (hsp_threshold_for). log_survivors: keep every window-filter
 * survivor as (query, subject, qp, sp, shape) in visit order. Returns a
 * handle for the accessors below. */
void* or_scan(const uint8_t* qres, const uint64_t* qoff, uint32_t nq, const int8_t* qbias,
              const uint8_t* sres, const uint64_t* soff, uint32_t ns, const int32_t* mat,
              const int32_t* region_thr, const or_scan_params* p, int log_survivors,
              or_scan_counters* ctr);
uint64_t or_scan_nregions(void* h, uint32_t q);
void or_scan_regions(void* h, uint32_t q, or_region* out);
uint64_t or_scan_nsurvivors(void* h);
void or_scan_survivors(void* h, uint32_t* out);
void or_scan_free(void* h);

#ifdef __cplusplus
}
#endif
#endif
