/*
 * saber_hip.h -- C ABI of the MI355X (gfx950) candidate-scoring library.
 *
 * This is the boundary a SABER host (the reference is a Rust crate) binds with
 * an `extern "C"` block: plain pointers and sizes, caller-owned outputs, int
 * status codes. It replaces, one entry point each:
 *
 *   saber_hip_score_windows  <- trait ScoreBackend::score_windows
 *                               (src/gpu/mod.rs:45-59; CUDA impl
 *                               src/gpu/cuda.rs:115-155, CPU twin
 *                               src/gpu/cpu_fallback.rs:22-44)
 *   saber_hip_score_bands    <- score_banded_many (src/algorithm/band_simd8.rs:78-185)
 *                               and banded_score (src/algorithm/banded.rs:408-449),
 *                               called from Engine::score_windows
 *                               (src/search/engine.rs:1055, 1081)
 *   saber_hip_score_full     <- score_many (src/algorithm/sw_simd.rs:95-153) with
 *                               the end cell of sw_reference
 *                               (src/algorithm/reference.rs:15-52)
 *   saber_hip_xdrop_batch    <- xdrop_cat / xdrop_avx2 / ungapped_xdrop_prof
 *                               (src/search/engine/scan.rs:859-904, 1024-1031;
 *                               src/search/engine.rs:2425-2455), called from
 *                               visit_scan (scan.rs:2414-2420)
 *
 * Conventions (src/gpu/mod.rs:9-31): residues are alignment codes 0..23
 * (ARNDCQEGHILKMFPSTWYVBZX*), any code >= 24 is scored as X (22); a gap of
 * length k costs gap_open + (k-1)*gap_extend; bias[i], when given, is added to
 * every aligned column of query position i. Every score is the exact int32
 * value of the reference CPU path. A call that cannot deliver exact results
 * returns a nonzero status (the host then falls back, as
 * Engine::score_windows_simd does on a backend error, engine.rs:1107-1116);
 * saber_hip_last_error() describes the last failure of the calling thread.
 *
 * Threading: a context serialises its own launches (like the Mutex in
 * src/gpu/cuda.rs:37, 144), so one context may be shared by threads.
 * Device memory is owned by the context and its batches; outputs are caller
 * owned host buffers.
 */
#ifndef SABER_HIP_H
#define SABER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SABER_HIP_OK 0
#define SABER_HIP_EINVAL 1   /* bad argument (lengths, indices, null pointers) */
#define SABER_HIP_EDEVICE 2  /* HIP runtime / device failure */
#define SABER_HIP_ENOMEM 3   /* device allocation failed */
#define SABER_HIP_ESTATE 4   /* call out of order (e.g. read before run) */

/* Context options. */
#define SABER_HIP_FORCE_EXACT32 1u /* score every band job with the int32 kernel (testing) */
#define SABER_HIP_NO_GRAPH 2u      /* enqueue band and full passes directly instead of replaying a captured graph */
#define SABER_HIP_INJECT_FAULTS 4u /* every scoring entry point (score_windows, score_bands, score_full, xdrop,
                                      band_align, traceback and the batch creators) fails with SABER_HIP_EDEVICE "injected fault" and writes
                                      no result: exercises a host's fallback path (backend_fallbacks,
                                      src/search/engine.rs:1107-1118) */
#define SABER_HIP_HOST_PLAN 8u   /* plan every band batch on the host (testing; batches of 65,536 jobs and more are
                                    otherwise planned on the device) */
#define SABER_HIP_DEVICE_PLAN 16u /* plan every band batch on the device, small ones too (testing) */
typedef struct saber_hip_opts {
    uint32_t flags;
    uint32_t reserved[7];
} saber_hip_opts;

/* Substitution matrix (24x24, row-major by query code) and SABER gap costs
 * (GapCosts, src/scoring/stats.rs:59-103). */
typedef struct saber_hip_scheme {
    int32_t matrix[576];
    int32_t gap_open;
    int32_t gap_extend;
} saber_hip_scheme;

/* A set of sequences: residues concatenated, offsets[count] + 1 entries.
 * bias (queries only) is NULL or holds one int8 per residue. */
typedef struct saber_hip_seqset {
    const uint8_t* residues;
    const uint64_t* offsets;
    uint32_t count;
    const int8_t* bias;
} saber_hip_seqset;

/* One banded job: query `query` against the window [win_lo, win_hi) of
 * subject `subject`, restricted to diagonals band_lo <= j - i <= band_hi in
 * window coordinates (BandJob, src/algorithm/band_simd8.rs:29-33; Band,
 * src/algorithm/banded.rs:30-54). */
typedef struct saber_hip_band_job {
    uint32_t query;
    uint32_t subject;
    uint32_t win_lo;
    uint32_t win_hi;
    int32_t band_lo;
    int32_t band_hi;
} saber_hip_band_job;

typedef struct saber_hip_ctx saber_hip_ctx;
typedef struct saber_hip_band_batch saber_hip_band_batch;

/* Library / context. device is a HIP device ordinal. */
const char* saber_hip_version(void);
const char* saber_hip_last_error(void);
int saber_hip_device_count(int* count);
int saber_hip_create(int device, const saber_hip_opts* opts, saber_hip_ctx** out);
void saber_hip_destroy(saber_hip_ctx* ctx);

/* Make a subject database device resident (the residues/subject_offsets of
 * SeedIndex, src/search/index.rs:644-703). Replaces any previous upload. */
int saber_hip_upload_db(saber_hip_ctx* ctx, const saber_hip_seqset* subjects);

/* A SABER index file, format v4 (src/search/index.rs:1-52), opened for its
 * residues and subject offsets: the sections `enc` and `enc_off` that
 * SeedIndex::residues() / subject_offsets() return (index.rs:644-706), read
 * through a read-only memory map. Open checks what SeedIndex::open
 * (index.rs:947-1009) and validate_structure (:822-849) check of them --
 * magic, version, byte order, header and metadata CRC-32, file size,
 * section bounds and alignment, the offsets against the residue section --
 * with the same messages; verify_sections also checks every section's
 * CRC-32 (verify_checksums, :925-934). Postings are not read. */
typedef struct saber_hip_index {
    void* mapping;                   /* internal */
    const uint8_t* residues;         /* total_residues alignment codes, every subject concatenated */
    const uint64_t* offsets;         /* num_subjects + 1 subject starts */
    uint64_t num_subjects, total_residues;
    int32_t protein;                 /* kind: 1 Protein, 0 Nucleotide */
    int32_t translated;              /* six-frame translated nucleotide database */
} saber_hip_index;
int saber_hip_index_open(const char* path, int verify_sections, saber_hip_index** out);
void saber_hip_index_close(saber_hip_index* idx);
/* saber_hip_upload_db of an opened index's residues and offsets. */
int saber_hip_upload_index(saber_hip_ctx* ctx, const saber_hip_index* idx);
/* CRC-32 of the index file (src/search/crc32.rs): IEEE, reflected, initial
 * and final xor 0xFFFFFFFF. */
uint32_t saber_hip_crc32(const uint8_t* data, uint64_t n);

/* ScoreBackend::score_windows: exact full-window SW scores of one query
 * against n windows (windows[k] has lens[k] residues), in input order.
 * bias may be NULL. An empty query gives all zeros. */
int saber_hip_score_windows(saber_hip_ctx* ctx, const uint8_t* query, uint32_t qlen,
                            const int8_t* bias, const uint8_t* const* windows,
                            const uint32_t* lens, uint32_t n,
                            const saber_hip_scheme* scheme, int32_t* out);

/* Banded scores with banded_score's end cell and cell count: per job the best
 * score, exclusive ends (q_end, s_end) of the first row-major optimal cell in
 * window coordinates (0, 0 when the score is 0) and the band cells inside the
 * rectangle. subjects == NULL selects the uploaded database. Any of q_end,
 * s_end, cells may be NULL; with both ends NULL only scores are computed
 * (score_banded_many's output, and faster). */
int saber_hip_score_bands(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                          const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                          uint64_t njobs, const saber_hip_scheme* scheme, int32_t* score,
                          uint32_t* q_end, uint32_t* s_end, uint64_t* cells);

/* Device-resident batch: create uploads the inputs and plans the work once;
 * run enqueues the whole scoring pass (asynchronous, on the context stream;
 * may be repeated); wait blocks until it is done; read copies the results.
 * A pass runs entirely on the device (output clear, the persistent packed
 * kernel, the int32 kernel over wide jobs and over the packed kernel's
 * saturated jobs, whose count it reads on the device); the first run is
 * enqueued directly, the second captures the pass as a HIP graph and later
 * runs replay it (SABER_HIP_NO_GRAPH in the context flags enqueues every run
 * directly).
 * ends = 0 computes scores only; reading end cells from such a batch is
 * SABER_HIP_ESTATE. */
int saber_hip_band_batch_create(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                                const saber_hip_seqset* subjects,
                                const saber_hip_band_job* jobs, uint64_t njobs,
                                const saber_hip_scheme* scheme, int ends, saber_hip_band_batch** out);
int saber_hip_band_batch_run(saber_hip_band_batch* b);
int saber_hip_band_batch_wait(saber_hip_band_batch* b);
int saber_hip_band_batch_read(saber_hip_band_batch* b, int32_t* score, uint32_t* q_end,
                              uint32_t* s_end, uint64_t* cells);
void saber_hip_band_batch_destroy(saber_hip_band_batch* b);

/* Timing of a batch: warmup + iters runs bracketed by HIP events on the
 * context stream. ms_run = mean wall time of one run on the device (as run
 * enqueues it: the graph replay); band_ms = mean time of the packed kernel's
 * launch (one launch per run, all shapes; measured over iters further passes
 * enqueued directly with events around the launch); variant_cells[v] = band
 * cells scored by variant v (v indexes saber_hip_band_variant); nvar is the
 * capacity of variant_cells on input and the number of variants on output.
 * cells = band cells per run. */
int saber_hip_band_batch_time(saber_hip_band_batch* b, int warmup, int iters, double* ms_run,
                              double* band_ms, uint64_t* variant_cells, int* nvar,
                              uint64_t* cells);

/* The plan a batch runs (no counterpart in the reference: the planner is
 * this library's own; tests compare the host and device planners with it).
 * npairs / nchunks receive the sizes; pairs (2 * npairs job ids, 0xFFFFFFFF
 * for none), blocks (4 * nchunks: shape, first pair, pairs, flags, in
 * dispatch order) and geometry (5 per job: clipped band lo, width, first row,
 * rows, transposed) are filled when not null. */
int saber_hip_band_batch_plan(saber_hip_band_batch* b, uint64_t* npairs, uint64_t* nchunks, uint32_t* pairs,
                              uint32_t* blocks, int32_t* geometry);

/* Banded traceback: banded_align_until (src/algorithm/banded.rs:472-521)
 * for each job -- the alignment of the job's best band cell, found by the
 * scalar recurrence of banded_align_scalar (banded.rs:227-333), the row stop of
 * the AVX2 kernels the reference dispatches to (banded_simd.rs:324, 594: rows
 * after the first holding band cells whose running best reaches stop_at[j]
 * are not computed; stop_at == NULL: none) and walk (banded.rs:336-403);
 * Engine::traceback (src/search/engine.rs:1328-1393) passes the window's exact
 * score as stop_at. Status per job: 0 alignment (LocalAlignment), 1 none
 * (Ok(None): empty input or band, best <= 0), 2 the direction matrix (query
 * length x clipped band width bytes) exceeds max_dir_bytes (Err). Query
 * coordinates and window coordinates (subject position - win_lo) are
 * half-open; the run-length operations ('M', 'I', 'D' with lengths, in
 * alignment order) of job j are the out[j].nops entries at out[j].op_off:
 * the jobs' operations are packed one after another (in no particular job
 * order), and op_codes / op_lens need room for the sum over all jobs of
 * query + window lengths (op_capacity), the most they can take. Runs on the
 * context stream; blocking.
 * Jobs run in launches whose device scratch (direction bytes, work rows,
 * operation bytes) fits a budget: 2048 MiB, or the environment variable
 * SABER_HIP_ALIGN_SCRATCH_MB (read at each call). */
typedef struct saber_hip_alignment {
    int32_t status;
    int32_t score;
    uint32_t q_start, q_end;
    uint32_t s_start, s_end;
    uint32_t nops;
    uint32_t reserved;
    uint64_t op_off;
    uint64_t cells;
} saber_hip_alignment;
int saber_hip_band_align(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                         const saber_hip_band_job* jobs, const int32_t* stop_at, uint64_t njobs,
                         const saber_hip_scheme* scheme, uint64_t max_dir_bytes, saber_hip_alignment* out,
                         uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity);
/* Engine::traceback (src/search/engine.rs:1328-1393) and its fallback
 * traceback_from_end (:1403-1446) for a batch of reported hits: per job the
 * banded traceback of the window, over the region's diagonals widened by a
 * pad (max(band_pad, 4), doubled after each attempt that neither reproduces
 * want nor covers the window; the full window when exhaustive or best == 0),
 * bounded by the scoring band when has_limit; a job whose band exceeds
 * max_dir_bytes locates the window's first optimal end cell with a score pass
 * over the scoring band (banded_score) and widens around that cell's
 * diagonal up to max(end score, want). The attempts of all jobs in a round
 * go to one saber_hip_band_align call. out[j] is the alignment (status 0)
 * or none (1); exact[j] = 1 when its score reaches want. Operations as for
 * saber_hip_band_align (packed, job j's out[j].nops entries at
 * out[j].op_off; capacity the sum of query + window lengths). counters may
 * be NULL; else it receives the engine's counters summed over the batch.
 * Runs on the context stream; blocking. */
typedef struct saber_hip_traceback_job {
    uint32_t query, subject;
    uint32_t win_lo, win_hi;     /* the window: subject residues [win_lo, win_hi) */
    int64_t d_min, d_max;        /* the region's diagonals (subject - query position) */
    int32_t best;                /* the region's best X-drop score (0: full window) */
    int32_t want;                /* the window's exact score */
    int32_t has_limit;           /* bound by the scoring band [limit_lo, limit_hi] */
    int32_t reserved;
    int64_t limit_lo, limit_hi;  /* window diagonals */
} saber_hip_traceback_job;
typedef struct saber_hip_traceback_counters {
    uint64_t tracebacks;              /* banded_align_until calls */
    uint64_t band_widenings;
    uint64_t band_capped;
    uint64_t end_anchored_tracebacks;
    uint64_t banded_cells;
} saber_hip_traceback_counters;
int saber_hip_traceback(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                        const saber_hip_traceback_job* jobs, uint64_t njobs, const saber_hip_scheme* scheme,
                        int64_t band_pad, int exhaustive, uint64_t max_dir_bytes, saber_hip_alignment* out,
                        uint8_t* exact, uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity,
                        saber_hip_traceback_counters* counters);
/* saber_hip_traceback with the counters also per job (job_counters[j], may be
 * NULL): a caller that replays the engine's reporting loop
 * (engine.rs:1173-1208: max_target_seqs, max_hsps, overlapping hits) over
 * speculatively traced hits sums the counters of the hits it used only.
 * Both entry points fail with SABER_HIP_ESTATE when subjects == NULL and the
 * database is replaced by saber_hip_upload_db during the call. */
int saber_hip_traceback_ex(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                           const saber_hip_traceback_job* jobs, uint64_t njobs, const saber_hip_scheme* scheme,
                           int64_t band_pad, int exhaustive, uint64_t max_dir_bytes, saber_hip_alignment* out,
                           uint8_t* exact, uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity,
                           saber_hip_traceback_counters* counters, saber_hip_traceback_counters* job_counters);
/* saber_hip_band_align, also timing iters repetitions of each kernel launch
 * with HIP events: *kernel_ms = their mean summed over the launches. */
int saber_hip_band_align_time(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                              const saber_hip_band_job* jobs, const int32_t* stop_at, uint64_t njobs,
                              const saber_hip_scheme* scheme, uint64_t max_dir_bytes, saber_hip_alignment* out,
                              uint8_t* op_codes, uint32_t* op_lens, uint64_t op_capacity, int iters,
                              double* kernel_ms);

/* Exhaustive mode (score_many over whole subjects, src/algorithm/sw_simd.rs:95-153,
 * with sw_reference's end cell, src/algorithm/reference.rs:15-52; the
 * exhaustive preset, src/search/presets.rs:318-331): every query against
 * every subject. Outputs are row-major [query][subject]: score, exclusive
 * ends of the first row-major optimal cell (0, 0 when the score is 0).
 * subjects == NULL selects the uploaded database. q_end / s_end may be NULL:
 * when both are, only scores are computed (score_many's output, and faster).
 * A batch computes end cells when created with ends != 0; reading ends from
 * a scores-only batch is SABER_HIP_ESTATE.
 * A full batch's pass runs on the device like a band batch's: one persistent
 * kernel per register class of the subject-length shapes drains a chunk list
 * of (shape, query) segments, longest queries first; pairs with subjects
 * longer than the widest shape, and pairs whose packed score saturates, are
 * rescored by the int32 kernel, their count read on the device. Runs after
 * the first replay a captured HIP graph (see the band batch). */
typedef struct saber_hip_full_batch saber_hip_full_batch;
int saber_hip_score_full(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                         const saber_hip_seqset* subjects, const saber_hip_scheme* scheme,
                         int32_t* score, uint32_t* q_end, uint32_t* s_end);
int saber_hip_full_batch_create(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                                const saber_hip_seqset* subjects, const saber_hip_scheme* scheme, int ends,
                                saber_hip_full_batch** out);
int saber_hip_full_batch_run(saber_hip_full_batch* b);
int saber_hip_full_batch_read(saber_hip_full_batch* b, int32_t* score, uint32_t* q_end,
                              uint32_t* s_end);
/* Mean ms per run (HIP events on the context stream), cells per run
 * (sum of query lengths x sum of subject lengths) and the pairs the last run
 * scored with the exact int32 kernel. */
int saber_hip_full_batch_time(saber_hip_full_batch* b, int warmup, int iters, double* ms_run,
                              uint64_t* cells, uint64_t* rescued);
void saber_hip_full_batch_destroy(saber_hip_full_batch* b);

/* One seed hit for X-drop extension: query position qp, subject position sp
 * (qp < query length, sp < subject length), as visit_scan passes them to
 * xdrop_cat (src/search/engine/scan.rs:2414-2420). */
typedef struct saber_hip_xdrop_hit {
    uint32_t query;
    uint32_t subject;
    uint32_t qp;
    uint32_t sp;
} saber_hip_xdrop_hit;

typedef struct saber_hip_xdrop_batch saber_hip_xdrop_batch;

/* Ungapped X-drop of every hit (xdrop_cat, scan.rs:859-904): score, the
 * half-open query range [q_lo, q_hi) and the steps taken (right steps incl. the
 * seed pair + left steps). The matrix of `scheme` is used (gap costs are not).
 * subjects == NULL selects the uploaded database; q_lo, q_hi, steps may be NULL. */
int saber_hip_xdrop(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                    const saber_hip_seqset* subjects, const saber_hip_xdrop_hit* hits,
                    uint64_t nhits, const saber_hip_scheme* scheme, int32_t x, int32_t* score,
                    uint32_t* q_lo, uint32_t* q_hi, uint32_t* steps);

/* Device-resident X-drop batch (same pattern as the band batch). time():
 * mean ms per run of the extension kernel (HIP events around it) and the
 * total steps per run. */
int saber_hip_xdrop_batch_create(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                                 const saber_hip_seqset* subjects, const saber_hip_xdrop_hit* hits,
                                 uint64_t nhits, const saber_hip_scheme* scheme, int32_t x,
                                 saber_hip_xdrop_batch** out);
int saber_hip_xdrop_batch_run(saber_hip_xdrop_batch* b);
int saber_hip_xdrop_batch_read(saber_hip_xdrop_batch* b, int32_t* score, uint32_t* q_lo,
                               uint32_t* q_hi, uint32_t* steps);
int saber_hip_xdrop_batch_time(saber_hip_xdrop_batch* b, int warmup, int iters, double* ms_run,
                               uint64_t* steps);
void saber_hip_xdrop_batch_destroy(saber_hip_xdrop_batch* b);
/* No counterpart in the reference: saber_hip_xdrop for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid device allocation). Every input and the output are device
 * pointers: residues as alignment codes (query and subject buffers, and the
 * bias when given, readable 16 bytes before the first residue and 64 after
 * the last, as the library's own uploads are), u64 offsets, hits as 4 u32
 * (query, subject, qp, sp; not range-checked: they are device data), out as
 * 4 u32 per hit (score, q_lo, q_hi, steps). Its scratch (claim counters, the
 * score table, 32 bytes per hit) is a fixed array in device memory, one per
 * GPU, held until the process ends: at most SABER_HIP_XDROP_NOALLOC_MAX_HITS
 * hits per call (EINVAL above). Calls on one GPU share that array: calls
 * that overlap on one GPU race. Runs on the context stream; blocking. */
#define SABER_HIP_XDROP_NOALLOC_MAX_HITS (1u << 20)
int saber_hip_xdrop_noalloc(saber_hip_ctx* ctx, const uint8_t* qres, const uint64_t* qoff, const int8_t* qbias,
                            const uint8_t* sres, const uint64_t* soff, const uint32_t* hits, uint64_t nhits,
                            const saber_hip_scheme* scheme, int32_t x, uint32_t* out);

/* Kernel variant table of the packed band kernel: variant v uses L lanes per
 * job pair with C band cells per lane and covers clipped widths
 * L*C-K+1 .. L*C; bias = 1 for the build that adds a query bias.
 * Returns SABER_HIP_EINVAL when v is out of range. */
int saber_hip_band_variant(int v, int* L, int* C, int* K, int* bias);

/* Diagnostic (no counterpart in the reference): time a plain device-to-device
 * copy of `bytes` (16-byte vector loads/stores) iters times; ms = mean per
 * copy. Used to calibrate traffic counters and report copy bandwidth. */
int saber_hip_copy_probe(saber_hip_ctx* ctx, uint64_t bytes, int iters, double* ms);

/* Width -> variant choice used by the planner (the autotuner's output):
 * widths lo..hi (inclusive) use variant shape (L, C, K); L = 0 routes them to
 * the exact int32 kernel. End-cell batches follow the choice only for shapes
 * of at most 21 cells per lane (C <= 21); for wider ones they keep their own
 * modelled choice. */
int saber_hip_band_choose(saber_hip_ctx* ctx, int w_lo, int w_hi, int L, int C, int K);

/* Measured shape choice (no counterpart in the reference; the autotuner of
 * the band planner). The sample is the batch's first jobs in order, up to
 * target_cells band cells (0: 2e9), so every trial times the batch's own mix
 * of widths and pairings. The band widths the packed kernel serves are split
 * into classes of widths covered by the same shapes; starting from the
 * modelled choice, classes holding at least 1 % of the sample's cells are
 * visited in decreasing share, each covering shape is tried for the class,
 * and it is kept when the sample's packed launch (scores only, HIP events)
 * gets clearly faster: the median of three timings (each the mean of three
 * launches) lower than the incumbent's by more than 1 % and by more than the
 * larger of the two spreads (max - min). The result becomes the context's
 * width -> shape table for batches created afterwards. report (may be NULL)
 * receives up to nreport visited classes (reserved = 1 where the shape
 * changed; gcups = the sample's rate after tuning); nclasses = their count. */
typedef struct saber_hip_tune_entry {
    int32_t w_lo, w_hi;              /* class widths */
    int32_t L, C, K, reserved;       /* chosen shape */
    uint64_t cells;                  /* the class's band cells in the sample */
    double gcups;                    /* the sample's rate after tuning */
} saber_hip_tune_entry;
int saber_hip_band_autotune(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                            const saber_hip_seqset* subjects, const saber_hip_band_job* jobs,
                            uint64_t njobs, const saber_hip_scheme* scheme, uint64_t target_cells,
                            saber_hip_tune_entry* report, int nreport, int* nclasses);

/* Full-matrix shape autotuner (no counterpart in the reference): subjects
 * are grouped by their modelled shape (the narrowest with L*C >= length);
 * per group, in decreasing share of the residues, every shape covering the
 * group's longest subject up to twice its width is timed on a sample (whole
 * queries spread over the length order, up to target_cells cells against all
 * subjects; 0 = 5e10) and kept when the pass gets clearly faster (the rule
 * of saber_hip_band_autotune: median of three, beyond 1 % and the spread).
 * The choice applies to full batches of this flavour (ends) created
 * afterwards on the context. report entries: w_lo / w_hi = the group's
 * subject lengths, L, C = the chosen shape (K = 0; reserved = 1 when it is
 * not the modelled one), cells = the group's cells in the sample, gcups = the
 * sample's rate after tuning. */
int saber_hip_full_autotune(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                            const saber_hip_scheme* scheme, int ends, uint64_t target_cells,
                            saber_hip_tune_entry* report, int nreport, int* nclasses);

/* Planner tuning (no counterpart in the reference): band jobs with more than
 * latency_rows rows are planned for low per-row latency (short lanes, their
 * blocks dispatched first); 0 turns this off, -1 restores the default rule
 * (jobs whose rows alone would take a large share of the batch's estimated
 * time). Applies to batches created afterwards. */
int saber_hip_band_tune(saber_hip_ctx* ctx, int latency_rows);

/* ---- Host pipeline: statistics thresholds (host code, no device) ----
 * Karlin-Altschul parameters and per-query statistics, restating
 * src/scoring/stats.rs: the two-stage band cut (floor_lower_bound, used by
 * score_windows, src/search/engine.rs:1057-1068), the region floor
 * (min_score, Engine::prepare engine.rs:759-786), E-values and bit scores
 * (the reporting cut and the prefilter's labels). Optional bit-score floors
 * are passed as NaN for "none". */
typedef struct saber_hip_karlin {
    double lambda, k, log_k, h;
    double alpha_d_lambda, beta;     /* length adjustment (alpha / lambda, beta) */
    int32_t calibrated;              /* published gapped parameters */
    int32_t even_scores_only;
} saber_hip_karlin;
typedef struct saber_hip_query_stats {
    double lambda, k, log_k;
    int64_t length_adjustment;
    double eff_query_len, eff_db_len, search_space;
    int32_t even_scores_only;
    int32_t per_subject;             /* 0: database model, 1: per-subject model */
    int64_t query_len;
    double num_seqs, alpha_d_lambda, beta;
} saber_hip_query_stats;
/* gapped_protein (stats.rs:351-396): matrix 62 (BLOSUM62) or 45 (BLOSUM45),
 * gap costs in SABER convention (open = cost of the first gap residue). */
int saber_hip_stats_gapped_protein(int matrix, int32_t gap_open, int32_t gap_extend, int allow_uncalibrated,
                                   saber_hip_karlin* out);
/* compute_length_adjustment (stats.rs:458-521), NCBI BLAST_ComputeLengthAdjustment. */
int saber_hip_stats_length_adjustment(double k, double log_k, double alpha_d_lambda, double beta,
                                      int64_t query_length, int64_t db_length, int64_t db_num_seqs,
                                      int64_t* adj, int* converged);
/* SearchStatistics::for_query (stats.rs:599-636). */
int saber_hip_stats_for_query(const saber_hip_karlin* p, uint64_t db_length, uint64_t db_num_seqs,
                              int per_subject, uint64_t query_len, saber_hip_query_stats* out);
/* QueryStatistics methods (stats.rs:646-760). */
double saber_hip_stats_space_for(const saber_hip_query_stats* q, uint64_t subject_len);
double saber_hip_stats_space_lower_bound(const saber_hip_query_stats* q, uint64_t subject_len);
double saber_hip_stats_floor_lower_bound(const saber_hip_query_stats* q, double evalue_cut, double min_bits,
                                         uint64_t subject_len, int64_t db_floor);
double saber_hip_stats_evalue_in(const saber_hip_query_stats* q, int64_t raw, double space);
double saber_hip_stats_evalue_for(const saber_hip_query_stats* q, int64_t raw, uint64_t subject_len);
double saber_hip_stats_evalue(const saber_hip_query_stats* q, int64_t raw);
double saber_hip_stats_bitscore(const saber_hip_query_stats* q, int64_t raw);
int64_t saber_hip_stats_min_score(const saber_hip_query_stats* q, double evalue_cut, double min_bits);
/* evalue[i] = evalue_for(raw[i], subject_len[i]) of a query of query_len[i]
 * residues (queries' statistics are recomputed when query_len changes). */
int saber_hip_stats_evalue_batch(const saber_hip_karlin* p, uint64_t db_length, uint64_t db_num_seqs,
                                 int per_subject, uint64_t n, const uint64_t* query_len,
                                 const uint64_t* subject_len, const int64_t* raw, double* evalue);

/* ---- Host pipeline: hits -> HSPs -> regions -> plans -> scores ---- */
typedef struct saber_hip_hsp {
    int64_t diag;                    /* sp - qp */
    uint32_t q_lo, q_hi;             /* query range of the extension (half-open) */
    int32_t score;
    uint32_t reserved;
} saber_hip_hsp;
typedef struct saber_hip_region {
    uint64_t subject;
    int64_t d_min, d_max;            /* diagonal range */
    int32_t best;                    /* best HSP score */
    uint32_t n_hsp;                  /* HSP evidence: count and saturating score sum */
    int32_t hsp_sum;
    uint32_t reserved;
} saber_hip_region;
typedef struct saber_hip_plan {
    uint64_t lo, hi;                 /* subject window [lo, hi) */
    int64_t band_lo, band_hi;        /* band relative to the window (banded plans) */
    int32_t banded;
    uint32_t reserved;
} saber_hip_plan;
typedef struct saber_hip_plan_params {
    int64_t window_slack;            /* engine.rs:180 (64) */
    int64_t score_band;              /* engine.rs:181 */
    int32_t hsp_threshold;           /* engine.rs:171 (40) */
    int32_t exhaustive;
} saber_hip_plan_params;
/* visit_scan (src/search/engine/scan.rs:2384-2431) for one (query, subject)
 * group: hits (qp[k], sp[k]) in visit order, with the X-drop result of every
 * hit already computed (saber_hip_xdrop: extensions are independent, so
 * scoring all of them and replaying the coverage rule is exact). A hit is
 * skipped when qp < the q_hi of the last extension on its diagonal; an
 * extension with score >= hsp_threshold becomes an HSP. hsps must hold n. */
int saber_hip_visit_hits(uint32_t qlen, uint32_t slen, const uint32_t* qp, const uint32_t* sp, uint64_t n,
                         const int32_t* score, const uint32_t* q_lo, const uint32_t* q_hi, int32_t hsp_threshold,
                         saber_hip_hsp* hsps, uint64_t* nhsp, uint64_t* extended, uint64_t* avoided);
/* finish_subject (src/search/engine.rs:2350-2412): HSPs (sorted in place by
 * diagonal) joined into regions within 2 * window_slack diagonals, kept when
 * best >= region_thr or (>= 2 HSPs and sum >= region_thr) (only if
 * region_thr > hsp_threshold), ordered by (best desc, d_min asc) and cut to
 * max_regions_per_subject. out must hold n. */
int saber_hip_finish_subject(uint64_t subject, saber_hip_hsp* hsps, uint64_t n, int64_t window_slack,
                             int32_t hsp_threshold, int32_t region_thr, uint32_t max_regions_per_subject,
                             saber_hip_region* out, uint32_t* nout);
/* Per-query region cap (engine.rs:486-495): regions sorted by (best desc,
 * subject, d_min, d_max); the first min(n, cap) are kept. */
int saber_hip_cap_regions(saber_hip_region* regs, uint64_t n, uint64_t cap, uint64_t* nout);
/* window + plan_windows (src/search/engine.rs:1223-1324): one plan per
 * region (subject_len[i] = length of region i's subject). */
int saber_hip_plan_windows(const saber_hip_region* regs, uint64_t n, uint64_t qlen, const uint64_t* subject_len,
                           const saber_hip_plan_params* p, saber_hip_plan* out);
/* Banded part of Engine::score_windows (src/search/engine.rs:1021-1099) over
 * the regions of many queries in two GPU passes: region i of query qidx[i]
 * with plan plans[i]. When 0 < narrow_band < full_band, banded plans are
 * first scored in the narrow band [d_min - narrow_band, d_max + narrow_band]
 * and only those reaching thr[i] (score_prefilter_frac x floor_lower_bound,
 * from the statistics functions) are rescored in the plan's band
 * (rescored[i] = 1); otherwise banded plans use the plan's band directly.
 * Unbanded plans are scored over the whole query x window rectangle.
 * subjects NULL = the uploaded database. */
int saber_hip_score_regions(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seqset* subjects,
                            const uint32_t* qidx, const saber_hip_region* regs, const saber_hip_plan* plans,
                            uint64_t n, const saber_hip_scheme* scheme, int64_t narrow_band, int64_t full_band,
                            const double* thr, int32_t* out, uint8_t* rescored);


/* ---- Search: the scan schedule end to end (Engine::search_scan,
 * src/search/engine/scan.rs:1261-1497) over the uploaded database ----
 *
 * Per query batch: preparation (composition bias, statistics, floors,
 * Engine::prepare engine.rs:759-786, masking off), seed tables of every shape
 * (build_batch scan.rs:1500-1596), the database scan with the 32-residue
 * window identity filter (scan_hits / filter_avx2 scan.rs:1603-1804,
 * 2242-2323), visits in (query position, shape, subject position) order per
 * (query, subject) with diagonal coverage and X-drop (visit_chunk_inner /
 * visit_scan scan.rs:2084-2225, 2384-2431), regions (finish_subject_into
 * scan.rs:2437-2474), the per-query cap, plan_windows, the two-stage banded
 * scores (score_windows engine.rs:1021-1099) and finish (E-values, the
 * progressive traceback, make_hit, engine.rs:1127-1221, 1447-1530).
 * Seed lookup, window filter, survivor ordering, X-drop, coverage and region
 * building run on the device; the per-query cap, planning, thresholds,
 * E-values and the reporting loop run on the host, the scoring and traceback
 * passes on the device. */
#define SABER_HIP_MAX_SHAPES 32
#define SABER_HIP_MAX_SHAPE_SPAN 64
typedef struct saber_hip_seed_scheme {
    uint8_t alphabet_map[24];          /* SeedAlphabet::map: reduced letter per code, 255 = not seedable */
    uint32_t alphabet_size;            /* SeedAlphabet::size */
    uint32_t nshapes;
    char shapes[SABER_HIP_MAX_SHAPES][SABER_HIP_MAX_SHAPE_SPAN + 4];  /* SeedShape patterns ("11110111"), NUL-terminated */
} saber_hip_seed_scheme;
/* SearchParams (engine.rs:62-197) fields on the path; saber_hip_search_defaults
 * fills the default preset (presets.rs:129-140) with the CLI defaults. */
typedef struct saber_hip_search_params {
    double evalue;                     /* -E: 10 */
    uint32_t max_target_seqs;          /* -k: 10 */
    uint32_t max_hsps;                 /* 1 */
    int32_t hsp_threshold;             /* 40 */
    int32_t x_drop_ungapped;           /* 20 */
    double hsp_floor_fraction;         /* 0.65 */
    uint32_t hit_filter_min_ids;       /* 27 (0 = no window filter) */
    uint32_t hit_filter_exact_weight;  /* 2 */
    int64_t band_pad;                  /* 16 */
    int64_t window_slack;              /* 64 */
    int64_t score_band;                /* 48 (0 = window scoring) */
    int64_t score_prefilter_band;      /* 16 (0 = one pass) */
    double score_prefilter_frac;       /* 0.75 */
    uint64_t max_candidates;           /* 5000 */
    uint32_t max_regions_per_subject;  /* 4 (at most 8 here) */
    int32_t comp_bias;                 /* 1 */
    uint64_t seed_max_occ;             /* 0 = auto_occ_cap */
    uint64_t traceback_max_bytes;      /* 32 MiB */
    uint64_t ref_db_residues;          /* 40 000 000 */
    double min_identity, min_query_cover, min_subject_cover;   /* 0 */
    int32_t per_subject_evalues;       /* 1: EvalueModel::PerSubject (default), 0: Database */
    uint32_t keep_regions;             /* 0 (default): only the hits; 1: also every capped region's record,
                                          for saber_hip_search_regions / saber_hip_search_features */
    uint32_t reserved[6];
} saber_hip_search_params;
void saber_hip_search_defaults(saber_hip_search_params* p);
/* The seed scheme of the default preset: Murphy-10 and DIAMOND's 16 weight-7
 * spaced seeds (presets.rs:41-58, alphabet.rs:325-327). */
void saber_hip_seed_scheme_default(saber_hip_seed_scheme* s);

/* A reported hit (Hit, engine.rs:312-337): 1-based starts, inclusive ends;
 * its run-length operations ('M', 'I', 'D') are runs [op_off, op_off + nops)
 * of its query's operation list. */
typedef struct saber_hip_hit {
    uint64_t subject;
    int32_t score;
    uint32_t exact_in_window;
    double bitscore, evalue;
    uint64_t q_start, q_end, s_start, s_end;
    uint64_t columns, matches, mismatches, positives, gap_opens, gap_columns;
    uint64_t q_len, s_len;
    uint64_t op_off;
    uint32_t nops, reserved;
} saber_hip_hit;
/* Counters of the run (Counters, engine.rs:203-297; the names are the
 * reference's --stats keys). */
typedef struct saber_hip_search_counters {
    uint64_t queries, query_residues, seed_positions, seeds_frequent_skipped;
    uint64_t scan_candidates, hits, hits_filtered, scan_groups;
    uint64_t ungapped_extensions, extensions_avoided, xdrop_steps, hsps;
    uint64_t regions, regions_dropped_by_cap, regions_weak_dropped, regions_below_evalue;
    uint64_t band_scored_regions, band_rescored, banded_cells;
    uint64_t tracebacks, band_widenings, band_capped, end_anchored_tracebacks;
    uint64_t filtered_identity_coverage, alignments;
    uint64_t regions_pruned;           /* removed by the learned prefilter (no reference counter) */
    uint64_t reserved[6];
} saber_hip_search_counters;
/* Wall time of each stage (ms, host clock around device-synchronised stages). */
typedef struct saber_hip_search_timing {
    double prepare, seed_tables, scan, sort, xdrop, regions, plan, prefilter, score, finish, total;
    double reserved[5];
} saber_hip_search_timing;
/* Learned prefilter between plan_windows and the scoring passes (the MLP of
 * saber_hip_mlp_*): regions whose logit is below threshold are not scored.
 * mean / inv_std standardise the features (may be NULL). */
typedef struct saber_hip_mlp saber_hip_mlp;
typedef struct saber_hip_prefilter {
    saber_hip_mlp* model;
    const float* mean;
    const float* inv_std;
    float threshold;
    uint32_t reserved;
} saber_hip_prefilter;
/* One candidate region of a query after the cap, with its plan and score
 * (score 0 when pruned); flags: 1 rescored in the full band, 2 its score
 * passes the E-value cut, 4 reported as a hit, 8 pruned by the prefilter. */
typedef struct saber_hip_search_region {
    uint64_t subject;
    int64_t d_min, d_max;
    int32_t best;
    uint32_t n_hsp;
    int32_t hsp_sum;
    int32_t score;
    uint64_t win_lo, win_hi;
    int64_t band_lo, band_hi;
    int32_t banded;
    uint32_t flags;
    double evalue;
} saber_hip_search_region;
typedef struct saber_hip_search saber_hip_search;
/* Runs the search of `queries` (alignment codes) against the uploaded
 * database. scoring: matrix and gap costs; karlin: gapped parameters
 * (saber_hip_stats_gapped_protein); prefilter may be NULL. The result holds
 * the hits, regions, counters and timings until saber_hip_search_destroy. */
int saber_hip_search_run(saber_hip_ctx* ctx, const saber_hip_seqset* queries, const saber_hip_seed_scheme* seeds,
                         const saber_hip_scheme* scoring, const saber_hip_karlin* karlin,
                         const saber_hip_search_params* params, const saber_hip_prefilter* prefilter,
                         saber_hip_search** out);
uint64_t saber_hip_search_hit_count(const saber_hip_search* s, uint32_t query);
int saber_hip_search_hits(const saber_hip_search* s, uint32_t query, saber_hip_hit* out);
/* The query's operation list (all its hits), op_count entries. */
uint64_t saber_hip_search_op_count(const saber_hip_search* s, uint32_t query);
int saber_hip_search_ops(const saber_hip_search* s, uint32_t query, uint8_t* codes, uint32_t* lens);
/* Every query's hits and operations in one call (host code; no counterpart
 * in the reference, for a caller that reads a whole result set): query q's
 * hits are out[hit_off[q] .. hit_off[q + 1]) and its operation list starts at
 * codes / lens + op_base[q] (the hits' op_off are relative to it); hit_off and
 * op_base hold nq + 1 offsets. Null arrays are skipped, so a first call with
 * the offsets alone sizes the others. */
int saber_hip_search_all_hits(const saber_hip_search* s, uint64_t* hit_off, saber_hip_hit* out, uint64_t* op_base,
                              uint8_t* codes, uint32_t* lens);
uint64_t saber_hip_search_region_count(const saber_hip_search* s, uint32_t query);
int saber_hip_search_regions(const saber_hip_search* s, uint32_t query, saber_hip_search_region* out);
/* Prefilter features (bf16 [n][64], the feature list of features.hip) of
 * every capped region of every query, in query order then region order;
 * mean / inv_std as for the prefilter. n = sum of region counts. */
int saber_hip_search_features(saber_hip_search* s, const float* mean, const float* inv_std, uint16_t* out);
int saber_hip_search_get_counters(const saber_hip_search* s, saber_hip_search_counters* out);
int saber_hip_search_get_timing(const saber_hip_search* s, saber_hip_search_timing* out);
void saber_hip_search_destroy(saber_hip_search* s);
/* composition_bias (src/scoring/bias.rs:80-122), host code. */
int saber_hip_composition_bias(const uint8_t* query, uint64_t n, const int32_t* matrix, int8_t* out);

/* ---- Learned region prefilter (no function in the reference) ----
 * The model the scope defines for the region prefilter (it plugs in between
 * the region cap and score_windows, src/search/engine/scan.rs:1476-1485 and
 * src/search/engine.rs:806-809): a 64 -> 256 -> 256 -> 1 MLP with ReLU,
 * bf16 weights and activations, fp32 accumulation (bf16 MFMA), trained with
 * the class-weighted logistic loss
 *     L = (1/n) sum_r w_r (softplus(y_r) - t_r y_r),  w_r = pos_weight if t_r = 1 else 1,
 * and Adam. Parameters are fp32 on the host side, in one array of
 * SABER_HIP_MLP_NPARAMS floats: W1[64][256], b1[256], W2[256][256], b2[256],
 * w3[256], b3 (h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 . w3 + b3).
 * Features are bf16 bit patterns (uint16), row-major [n][64]. */
/* Prefilter feature vectors (bf16 bit patterns, [n][64]) of candidate
 * regions on the device: region i is band job jobs[i] (its query, subject
 * window and band) with X-drop results xscore/xq_lo/xq_hi of its
 * hits_per_region (<= 64) hits hits[i * hits_per_region ..]; the feature list
 * is documented in saber_amd/csrc/features.hip (X-drop evidence, geometry,
 * query and window composition). mean / inv_std (64 each, or both NULL)
 * standardise each feature, (f - mean) * inv_std, before the bf16 rounding. */
int saber_hip_region_features(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                              const saber_hip_seqset* subjects, const saber_hip_band_job* jobs, uint64_t n,
                              const saber_hip_xdrop_hit* hits, const int32_t* xscore, const uint32_t* xq_lo,
                              const uint32_t* xq_hi, uint32_t hits_per_region, int32_t hsp_threshold,
                              const float* mean, const float* inv_std, uint16_t* features);
/* The same, also timing the feature kernel (a diagnostic for the prefilter's
 * cost): kernel_ms = mean of iters further launches on the uploaded inputs,
 * HIP events on the context stream. */
int saber_hip_region_features_time(saber_hip_ctx* ctx, const saber_hip_seqset* queries,
                                   const saber_hip_seqset* subjects, const saber_hip_band_job* jobs, uint64_t n,
                                   const saber_hip_xdrop_hit* hits, const int32_t* xscore, const uint32_t* xq_lo,
                                   const uint32_t* xq_hi, uint32_t hits_per_region, int32_t hsp_threshold,
                                   const float* mean, const float* inv_std, uint16_t* features, int iters,
                                   double* kernel_ms);
#define SABER_HIP_MLP_IN 64
#define SABER_HIP_MLP_HID 256
#define SABER_HIP_MLP_NPARAMS (64 * 256 + 256 + 256 * 256 + 256 + 256 + 1)
/* params may be NULL: He-uniform initialisation from `seed`. */
int saber_hip_mlp_create(saber_hip_ctx* ctx, const float* params, uint64_t seed, saber_hip_mlp** out);
int saber_hip_mlp_get_params(saber_hip_mlp* m, float* params);
int saber_hip_mlp_set_params(saber_hip_mlp* m, const float* params);   /* also resets Adam state */
/* logits[n] of host features x[n][64]. */
int saber_hip_mlp_forward(saber_hip_mlp* m, const uint16_t* x, uint64_t n, float* logits);
/* Gradient of L over the batch (same layout as the parameters) and L itself;
 * parameters unchanged. */
int saber_hip_mlp_backward(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n,
                           float pos_weight, float* grads, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; loss may be NULL. */
int saber_hip_mlp_train_step(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n,
                             float pos_weight, float lr, float* loss);
/* Device-resident data set: upload once, then train / score slices of it
 * without host traffic. */
int saber_hip_mlp_upload(saber_hip_mlp* m, const uint16_t* x, const float* labels, uint64_t n);
int saber_hip_mlp_train_resident(saber_hip_mlp* m, uint64_t row0, uint64_t n, float pos_weight,
                                 float lr, float* loss);
int saber_hip_mlp_forward_resident(saber_hip_mlp* m, uint64_t row0, uint64_t n, float* logits);
/* Timing (HIP events on the context stream) of forward and of one training
 * step over rows [0, n) of the resident data: mean ms per call. */
int saber_hip_mlp_time(saber_hip_mlp* m, uint64_t n, int warmup, int iters, double* fwd_ms,
                       double* step_ms);
void saber_hip_mlp_destroy(saber_hip_mlp* m);

#ifdef __cplusplus
}
#endif
#endif /* SABER_HIP_H */
