/*
 * AVX2 CPU baseline for batched banded SW -- TEST / BENCH INFRASTRUCTURE ONLY.
 *
 * A C restatement of the reference's AVX2 band-scoring path, timed on the host
 * cores beside the GPU (the reference is Rust and cannot be built here):
 *
 *  score_banded_many (src/algorithm/band_simd8.rs:78-185): bands clipped to
 *    their rectangle; the u8 kernel (32 jobs of one query per register, lane
 *    coordinates c = j - lo, unsigned saturating arithmetic with substitution
 *    offset, band_simd8.rs:236-388) scores every job of width <= 126 whose
 *    lane maximum stays below 255 - pmax; other jobs go to
 *  banded_score (src/algorithm/banded.rs:408-449): the int16 intra-sequence
 *    kernel (16 band cells per register, E resolved by a max-plus prefix scan
 *    with a carried chunk value, banded_simd.rs:393-599) which declines when a
 *    row maximum exceeds 32767 - 16*ge - 192, then the exact int32 kernel.
 *
 * Only scores are produced (score_banded_many's output). Every result is
 * checked against the scalar oracle in tests/test_cpu_baseline.py.
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define U8_MAX_WIDTH 126
#define INVALID 31
#define NEG16 (-31000)

static inline int code(uint8_t c) { return c < 24 ? c : 22; }

/* ---------------- u8 inter-sequence kernel (up to 32 jobs of one query) -- */

typedef struct {
    const uint8_t* s;
    int64_t n;
    int64_t lo, w; /* clipped band: lo, hi - lo */
} u8job;

/* Lane maxima of up to 32 banded jobs of query q. */
static void batch_u8(const uint8_t* q, int64_t m, const int8_t* bias, const u8job* jobs, int nl,
                     const int32_t* mat, int32_t go, int32_t ge, int32_t off, uint8_t* maxes,
                     __m256i* H, __m256i* E, __m256i* B, uint8_t* cols, int64_t cols_cap) {
    int64_t wmax = 0;
    for (int l = 0; l < nl; l++)
        if (jobs[l].w > wmax) wmax = jobs[l].w;
    const int64_t ncols = m + wmax;
    (void)cols_cap;
    /* column-major residues in lane coordinates, INVALID outside a subject */
    memset(cols, INVALID, (size_t)ncols * 32);
    for (int l = 0; l < nl; l++) {
        int64_t c0 = -jobs[l].lo > 0 ? -jobs[l].lo : 0;
        int64_t c1 = jobs[l].n - jobs[l].lo < ncols ? jobs[l].n - jobs[l].lo : ncols;
        for (int64_t c = c0; c < c1; c++) cols[c * 32 + l] = (uint8_t)code(jobs[l].s[c + jobs[l].lo]);
    }
    int8_t wp1[32];
    for (int l = 0; l < 32; l++) wp1[l] = l < nl ? (int8_t)(jobs[l].w + 1) : 0;
    const __m256i wp1v = _mm256_loadu_si256((const __m256i*)wp1);
    for (int64_t i = 0; i < m; i++) {
        H[i] = _mm256_setzero_si256();
        E[i] = _mm256_setzero_si256();
        B[i] = _mm256_set1_epi8(bias ? bias[i] : 0);
    }
    /* shuffle tables: S(a, x) + off for x in 0..15 and 16..23 */
    __m256i tlo[24], thi[24];
    for (int a = 0; a < 24; a++) {
        uint8_t t0[16], t1[16];
        for (int x = 0; x < 16; x++) t0[x] = (uint8_t)(mat[a * 24 + x] + off);
        for (int x = 0; x < 16; x++) t1[x] = x < 8 ? (uint8_t)(mat[a * 24 + 16 + x] + off) : 0;
        tlo[a] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)t0));
        thi[a] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)t1));
    }
    const __m256i gov = _mm256_set1_epi8((char)go), gev = _mm256_set1_epi8((char)ge);
    const __m256i offv = _mm256_set1_epi8((char)off), zero = _mm256_setzero_si256();
    const __m256i nib = _mm256_set1_epi8(0x0f), fifteen = _mm256_set1_epi8(15);
    const __m256i invalid = _mm256_set1_epi8(INVALID), one = _mm256_set1_epi8(1);
    __m256i max_h = zero, prof[24];
    uint8_t qc[4096];
    const uint8_t* qcodes = q;
    uint8_t* qtmp = NULL;
    if (m <= 4096) {
        for (int64_t i = 0; i < m; i++) qc[i] = (uint8_t)code(q[i]);
        qcodes = qc;
    } else {
        qtmp = (uint8_t*)malloc((size_t)m);
        for (int64_t i = 0; i < m; i++) qtmp[i] = (uint8_t)code(q[i]);
        qcodes = qtmp;
    }
    for (int64_t c = 0; c < ncols; c++) {
        int64_t r0 = c - wmax > 0 ? c - wmax : 0;
        int64_t r1 = c < m - 1 ? c : m - 1;
        if (r0 > r1) continue;
        const __m256i rv = _mm256_loadu_si256((const __m256i*)(cols + c * 32));
        const __m256i colv = _mm256_cmpgt_epi8(invalid, rv);
        const __m256i idx = _mm256_and_si256(rv, nib);
        const __m256i ishi = _mm256_cmpgt_epi8(rv, fifteen);
        for (int a = 0; a < 24; a++)
            prof[a] = _mm256_blendv_epi8(_mm256_shuffle_epi8(tlo[a], idx),
                                         _mm256_shuffle_epi8(thi[a], idx), ishi);
        __m256i ul_h = r0 > 0 ? H[r0 - 1] : zero, u_h = zero, u_f = zero;
        __m256i dv = _mm256_set1_epi8((char)(c - r0));
        for (int64_t i = r0; i <= r1; i++) {
            const __m256i valid = _mm256_and_si256(colv, _mm256_cmpgt_epi8(wp1v, dv));
            const __m256i p = _mm256_add_epi8(prof[qcodes[i]], B[i]);
            const __m256i ph = H[i];
            const __m256i e = _mm256_and_si256(
                _mm256_max_epu8(_mm256_subs_epu8(ph, gov), _mm256_subs_epu8(E[i], gev)), valid);
            const __m256i f = _mm256_and_si256(
                _mm256_max_epu8(_mm256_subs_epu8(u_h, gov), _mm256_subs_epu8(u_f, gev)), valid);
            const __m256i diag = _mm256_subs_epu8(_mm256_adds_epu8(ul_h, p), offv);
            const __m256i h = _mm256_and_si256(_mm256_max_epu8(_mm256_max_epu8(e, f), diag), valid);
            max_h = _mm256_max_epu8(max_h, h);
            ul_h = ph;
            u_h = h;
            u_f = f;
            H[i] = h;
            E[i] = e;
            dv = _mm256_sub_epi8(dv, one);
        }
    }
    _mm256_storeu_si256((__m256i*)maxes, max_h);
    free(qtmp);
}

/* ---------------- int16 intra-sequence banded kernel (score only) ------- */

static inline __m256i shift_in16(__m256i v, __m256i prev) {
    /* [prev[15], v0 .. v14] */
    __m256i t = _mm256_permute2x128_si256(prev, v, 0x21);
    return _mm256_alignr_epi8(v, t, 14);
}

static inline __m256i prefix_max16(__m256i v, __m256i negv) {
    v = _mm256_max_epi16(v, _mm256_alignr_epi8(v, negv, 14));
    v = _mm256_max_epi16(v, _mm256_alignr_epi8(v, negv, 12));
    v = _mm256_max_epi16(v, _mm256_alignr_epi8(v, negv, 8));
    __m256i t = _mm256_permute2x128_si256(v, v, 0x08);
    __m256i b = _mm256_shuffle_epi32(_mm256_shufflehi_epi16(t, 0xFF), 0xFF);
    return _mm256_max_epi16(v, _mm256_blend_epi32(negv, b, 0xF0));
}

static inline int hmax16(__m256i v) {
    __m128i a = _mm_max_epi16(_mm256_castsi256_si128(v), _mm256_extracti128_si256(v, 1));
    a = _mm_max_epi16(a, _mm_shuffle_epi32(a, 0x4E));
    a = _mm_max_epi16(a, _mm_shuffle_epi32(a, 0xB1));
    a = _mm_max_epi16(a, _mm_shufflelo_epi16(a, 0xB1));
    return (int16_t)_mm_extract_epi16(a, 0);
}

/* Returns 1 and the score when every value stayed within the int16 range,
 * 0 when the caller must rerun exactly (banded_simd.rs:412-415, 580-583). */
static int banded_i16(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                      const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge, int64_t lo,
                      int64_t w, int16_t* buf, int8_t* prof, int32_t* out) {
    if (go < 0 || go > 1000 || ge < 0 || ge > 1000 || ge > go) return 0;
    const int32_t limit = 32767 - 16 * ge - 192;
    const int64_t wp = w + 32;
    int16_t *hp = buf, *fp = buf + wp, *hc = buf + 2 * wp, *fc = buf + 3 * wp;
    for (int64_t k = 0; k < wp; k++) {
        hp[k] = hc[k] = 0;
        fp[k] = fc[k] = NEG16;
    }
    /* subject profile rows for the query letters: prof[a * stride + j] */
    const int64_t stride = (n + 32 + 15) & ~(int64_t)15;
    uint32_t present = 0;
    for (int64_t i = 0; i < m; i++) present |= 1u << code(q[i]);
    /* row 24: subject codes padded with '*' (23); rows a: 16-byte shuffles */
    int8_t* codes = prof + 24 * stride;
    for (int64_t j = 0; j < stride; j++) codes[j] = (int8_t)(j < n ? code(s[j]) : 23);
    const __m128i nib = _mm_set1_epi8(0x0f), fif = _mm_set1_epi8(15);
    for (int a = 0; a < 24; a++) {
        if (!(present >> a & 1)) continue;
        int8_t lo8[16], hi8[16];
        for (int x = 0; x < 16; x++) {
            int v = mat[a * 24 + x];
            lo8[x] = (int8_t)(v < -128 ? -128 : v > 127 ? 127 : v);
            v = x < 8 ? mat[a * 24 + 16 + x] : 0;
            hi8[x] = (int8_t)(v < -128 ? -128 : v > 127 ? 127 : v);
        }
        const __m128i tl = _mm_loadu_si128((const __m128i*)lo8), th = _mm_loadu_si128((const __m128i*)hi8);
        int8_t* row = prof + a * stride;
        for (int64_t j = 0; j < stride; j += 16) {
            const __m128i sv = _mm_loadu_si128((const __m128i*)(codes + j));
            const __m128i idx = _mm_and_si128(sv, nib);
            const __m128i v = _mm_blendv_epi8(_mm_shuffle_epi8(tl, idx), _mm_shuffle_epi8(th, idx),
                                              _mm_cmpgt_epi8(sv, fif));
            _mm_storeu_si128((__m128i*)(row + j), v);
        }
    }
    const __m256i zero = _mm256_setzero_si256(), negv = _mm256_set1_epi16(NEG16);
    const __m256i gov = _mm256_set1_epi16((short)go), gev = _mm256_set1_epi16((short)ge);
    const __m256i lane = _mm256_setr_epi16(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    const __m256i ramp = _mm256_mullo_epi16(lane, gev);
    const __m256i ge16 = _mm256_set1_epi16((short)(16 * ge));
    int32_t best = 0;
    for (int64_t i = 0; i < m; i++) {
        int64_t kmin = -(i + lo) > 0 ? -(i + lo) : 0;
        int64_t kmax = n - 1 - i - lo < w - 1 ? n - 1 - i - lo : w - 1;
        if (kmin > kmax) {
            for (int64_t k = 0; k < wp; k++) {
                hc[k] = k < kmin ? 0 : NEG16;
                fc[k] = NEG16;
            }
        } else {
            for (int64_t k = 0; k < kmin; k++) {
                hc[k] = 0;
                fc[k] = NEG16;
            }
            const int8_t* prow = prof + code(q[i]) * stride;
            const __m256i bv = _mm256_set1_epi16((short)(bias ? bias[i] : 0));
            __m256i carry = negv, rowmax = zero;
            const int64_t base = i + lo;
            for (int64_t k0 = kmin; k0 <= kmax; k0 += 16) {
                const int64_t nvalid = kmax + 1 - k0 < 16 ? kmax + 1 - k0 : 16;
                const __m256i sub = _mm256_cvtepi8_epi16(_mm_loadu_si128((const __m128i*)(prow + base + k0)));
                const __m256i hpk = _mm256_loadu_si256((const __m256i*)(hp + k0));
                const __m256i hp1 = _mm256_loadu_si256((const __m256i*)(hp + k0 + 1));
                const __m256i fp1 = _mm256_loadu_si256((const __m256i*)(fp + k0 + 1));
                const __m256i d = _mm256_adds_epi16(_mm256_adds_epi16(hpk, sub), bv);
                const __m256i f = _mm256_max_epi16(_mm256_subs_epi16(hp1, gov), _mm256_subs_epi16(fp1, gev));
                const __m256i hprime = _mm256_max_epi16(_mm256_max_epi16(d, f), zero);
                const __m256i valid = _mm256_cmpgt_epi16(_mm256_set1_epi16((short)nvalid), lane);
                const __m256i hprime_v = _mm256_blendv_epi8(negv, hprime, valid);
                const __m256i open = _mm256_subs_epi16(hprime_v, gov);
                const __m256i e_loc = _mm256_subs_epi16(
                    prefix_max16(_mm256_adds_epi16(shift_in16(open, negv), ramp), negv), ramp);
                const __m256i e = _mm256_max_epi16(e_loc, _mm256_subs_epi16(carry, ramp));
                const __m256i h = _mm256_max_epi16(hprime, e);
                const __m256i h_v = _mm256_blendv_epi8(negv, h, valid);
                const __m256i f_v = _mm256_blendv_epi8(negv, f, valid);
                _mm256_storeu_si256((__m256i*)(hc + k0), h_v);
                _mm256_storeu_si256((__m256i*)(fc + k0), f_v);
                rowmax = _mm256_max_epi16(rowmax, h_v);
                const __m256i a = _mm256_max_epi16(_mm256_subs_epi16(hprime_v, gov), _mm256_subs_epi16(e_loc, gev));
                const __m256i a15 = _mm256_shufflehi_epi16(
                    _mm256_shufflelo_epi16(_mm256_permute4x64_epi64(a, 0xFF), 0xFF), 0xFF);
                carry = _mm256_max_epi16(a15, _mm256_subs_epi16(carry, ge16));
            }
            const int32_t rm = hmax16(rowmax);
            if (rm > limit) return 0;
            if (rm > best) best = rm;
            _mm256_storeu_si256((__m256i*)(hc + kmax + 1), negv);
            _mm256_storeu_si256((__m256i*)(fc + kmax + 1), negv);
        }
        int16_t* t = hp;
        hp = hc;
        hc = t;
        t = fp;
        fp = fc;
        fc = t;
    }
    *out = best;
    return 1;
}

/* ---------------- score_banded_many for one query ----------------------- */

typedef struct {
    int64_t k, w, n;
} order_t;

static int order_cmp(const void* a, const void* b) {
    const order_t *x = (const order_t*)a, *y = (const order_t*)b;
    if (x->w != y->w) return x->w < y->w ? -1 : 1;
    if (x->n != y->n) return x->n < y->n ? -1 : 1;
    return x->k < y->k ? -1 : x->k > y->k;
}

typedef struct {
    __m256i *H, *E, *B;
    int64_t hcap;
    uint8_t* cols;
    int64_t ccap;
    int16_t* buf;
    int64_t bcap;
    int8_t* prof;
    int64_t pcap;
    order_t* ord;
    int64_t ocap;
} scratch_t;

static void* grow(void* p, int64_t* cap, int64_t need, size_t elem, size_t align) {
    if (need <= *cap) return p;
    free(p);
    *cap = need * 2;
    void* q = NULL;
    if (posix_memalign(&q, align, (size_t)*cap * elem)) return NULL;
    return q;
}

static int32_t exact_score(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                           const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge,
                           int64_t lo, int64_t w, scratch_t* sc) {
    sc->buf = (int16_t*)grow(sc->buf, &sc->bcap, 4 * (w + 32) + 16, 2, 32);
    sc->prof = (int8_t*)grow(sc->prof, &sc->pcap, 25 * (((n + 32 + 15) & ~15) + 16), 1, 32);
    int32_t r;
    if (banded_i16(q, m, s, n, bias, mat, go, ge, lo, w, sc->buf, sc->prof, &r)) return r;
    int64_t qe, se;
    uint64_t cells;
    or_banded_score(q, m, s, n, bias, mat, go, ge, lo, lo + w - 1, &r, &qe, &se, &cells);
    return r;
}

static void score_banded_many(const uint8_t* q, int64_t m, const int8_t* bias,
                              const uint8_t* const* subj, const int64_t* slen, const int64_t* blo,
                              const int64_t* bhi, int64_t njobs, const int32_t* mat, int32_t go,
                              int32_t ge, int32_t* out, scratch_t* sc) {
    for (int64_t k = 0; k < njobs; k++) out[k] = 0;
    if (njobs == 0 || m == 0) return;
    int smin = 1 << 30, smax = -(1 << 30), bmin = 0, bmax = 0;
    for (int t = 0; t < 576; t++) {
        if (mat[t] < smin) smin = mat[t];
        if (mat[t] > smax) smax = mat[t];
    }
    if (bias) {
        bmin = 127;
        bmax = -128;
        for (int64_t i = 0; i < m; i++) {
            if (bias[i] < bmin) bmin = bias[i];
            if (bias[i] > bmax) bmax = bias[i];
        }
    }
    int t0 = smin + (bmin < 0 ? bmin : 0);
    const int off = -(t0 < 0 ? t0 : 0);
    const int pmax = smax + (bmax > 0 ? bmax : 0) + off;
    const int gaps_ok = go >= 0 && go <= 255 && ge >= 0 && ge <= 255;
    const int u8_ok = pmax < 128 && gaps_ok && off <= 127;
    sc->ord = (order_t*)grow(sc->ord, &sc->ocap, njobs, sizeof(order_t), 16);
    int64_t nord = 0;
    for (int64_t k = 0; k < njobs; k++) {
        const int64_t n = slen[k];
        if (n == 0 || bhi[k] < blo[k]) continue;
        const int64_t lo = blo[k] > -(m - 1) ? blo[k] : -(m - 1);
        const int64_t hi = bhi[k] < n - 1 ? bhi[k] : n - 1;
        if (hi < lo) continue;
        if (u8_ok && hi - lo <= U8_MAX_WIDTH) {
            order_t o = {k, hi - lo, n};
            sc->ord[nord++] = o;
        } else {
            out[k] = exact_score(q, m, subj[k], n, bias, mat, go, ge, lo, hi - lo + 1, sc);
        }
    }
    if (nord == 0) return;
    qsort(sc->ord, (size_t)nord, sizeof(order_t), order_cmp);
    const int limit = 255 - pmax;
    sc->H = (__m256i*)grow(sc->H, &sc->hcap, 3 * m, sizeof(__m256i), 32);
    sc->E = sc->H + m;
    sc->B = sc->H + 2 * m;
    for (int64_t c = 0; c < nord; c += 32) {
        const int nl = nord - c < 32 ? (int)(nord - c) : 32;
        u8job jobs[32];
        int64_t wmax = 0;
        for (int l = 0; l < nl; l++) {
            const int64_t k = sc->ord[c + l].k;
            const int64_t n = slen[k];
            const int64_t lo = blo[k] > -(m - 1) ? blo[k] : -(m - 1);
            const int64_t hi = bhi[k] < n - 1 ? bhi[k] : n - 1;
            jobs[l].s = subj[k];
            jobs[l].n = n;
            jobs[l].lo = lo;
            jobs[l].w = hi - lo;
            if (hi - lo > wmax) wmax = hi - lo;
        }
        sc->cols = (uint8_t*)grow(sc->cols, &sc->ccap, 32 * (m + wmax + 1), 1, 32);
        uint8_t maxes[32];
        batch_u8(q, m, bias, jobs, nl, mat, go, ge, off, maxes, sc->H, sc->E, sc->B, sc->cols, sc->ccap);
        for (int l = 0; l < nl; l++) {
            const int64_t k = sc->ord[c + l].k;
            if (maxes[l] >= limit)
                out[k] = exact_score(q, m, jobs[l].s, jobs[l].n, bias, mat, go, ge, jobs[l].lo,
                                     jobs[l].w + 1, sc);
            else
                out[k] = maxes[l];
        }
    }
}

/* ---------------- batch driver ------------------------------------------ */

typedef struct {
    const uint8_t *qres, *sres;
    const int8_t* bias;
    const uint64_t* qoff;
    const uint32_t* qlen;
    const uint64_t* soff;
    const uint32_t* slen;
    const int32_t *lo, *hi;
    const int64_t* grp; /* group starts (jobs of one query are consecutive) */
    int64_t g0, g1;
    const int32_t* mat;
    int32_t go, ge;
    int32_t* score;
} drv_t;

static void* drv_worker(void* p) {
    drv_t* a = (drv_t*)p;
    scratch_t sc;
    memset(&sc, 0, sizeof(sc));
    int64_t cap = 0;
    const uint8_t** subj = NULL;
    int64_t *sl = NULL, *bl = NULL, *bh = NULL;
    for (int64_t g = a->g0; g < a->g1; g++) {
        const int64_t k0 = a->grp[g], k1 = a->grp[g + 1], nj = k1 - k0;
        if (nj > cap) {
            cap = nj * 2;
            subj = (const uint8_t**)realloc(subj, sizeof(*subj) * cap);
            sl = (int64_t*)realloc(sl, 8 * cap);
            bl = (int64_t*)realloc(bl, 8 * cap);
            bh = (int64_t*)realloc(bh, 8 * cap);
        }
        for (int64_t t = 0; t < nj; t++) {
            subj[t] = a->sres + a->soff[k0 + t];
            sl[t] = a->slen[k0 + t];
            bl[t] = a->lo[k0 + t];
            bh[t] = a->hi[k0 + t];
        }
        score_banded_many(a->qres + a->qoff[k0], a->qlen[k0], a->bias ? a->bias + a->qoff[k0] : NULL,
                          subj, sl, bl, bh, nj, a->mat, a->go, a->ge, a->score + k0, &sc);
    }
    free(subj);
    free(sl);
    free(bl);
    free(bh);
    free(sc.H);
    free(sc.cols);
    free(sc.buf);
    free(sc.prof);
    free(sc.ord);
    return NULL;
}

/* Scores of a band batch in the flat layout of or_banded_many. Consecutive
 * jobs with the same query offset form one score_banded_many call; groups are
 * split across nthreads threads. */
void cpu_banded_many_avx2(const uint8_t* qres, const uint8_t* sres, const int8_t* bias,
                          const uint64_t* qoff, const uint32_t* qlen, const uint64_t* soff,
                          const uint32_t* slen, const int32_t* lo, const int32_t* hi, int64_t njobs,
                          const int32_t* mat, int32_t go, int32_t ge, int32_t* score, int nthreads) {
    if (njobs == 0) return;
    int64_t* grp = (int64_t*)malloc(8 * (size_t)(njobs + 1));
    int64_t ng = 0;
    for (int64_t k = 0; k < njobs; k++)
        if (k == 0 || qoff[k] != qoff[k - 1] || qlen[k] != qlen[k - 1]) grp[ng++] = k;
    grp[ng] = njobs;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    drv_t args[256];
    for (int t = 0; t < nthreads; t++) {
        drv_t a = {qres, sres, bias, qoff, qlen, soff, slen, lo, hi, grp,
                   ng * t / nthreads, ng * (t + 1) / nthreads, mat, go, ge, score};
        args[t] = a;
        pthread_create(&th[t], NULL, drv_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(grp);
}

/* ---- score_many (src/algorithm/sw_simd.rs:95-153): full-matrix SW scores
 * of one query against many subjects. Subjects longest first (ties by
 * index); an 8-bit pass over 32 subjects per register (sw_simd.rs:287-427:
 * unsigned saturating lanes with the substitution offset, exact when the
 * lane maximum stays below 255 - pmax); its other subjects re-scored by the
 * 16-bit pass over 16 subjects per register (sw_simd.rs:157-277); a 16-bit
 * score reaching 32767 repaired by the int32 full-band score. ---- */

typedef struct {
    const uint8_t* s;
    int64_t n;
    int64_t idx;
} mjob;

static int mjob_cmp(const void* a, const void* b) {
    const mjob *x = (const mjob*)a, *y = (const mjob*)b;
    if (x->n != y->n) return x->n > y->n ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* 16-bit pass over the subjects of `jobs` (already in order) */
static void many_i16(const uint8_t* q, int64_t m, const int8_t* bias, const mjob* jobs, int64_t nj,
                     const int32_t* mat, int32_t go, int32_t ge, __m256i* H, __m256i* E, __m256i* B,
                     int32_t* out) {
    __m128i tlo[24], thi[24];
    for (int a = 0; a < 24; ++a) {
        int8_t lo[16], hi[16];
        memset(hi, 0, sizeof hi);
        for (int x = 0; x < 16; ++x) {
            int v = mat[a * 24 + x];
            lo[x] = (int8_t)(v < -128 ? -128 : v > 127 ? 127 : v);
        }
        for (int x = 16; x < 24; ++x) {
            int v = mat[a * 24 + x];
            hi[x - 16] = (int8_t)(v < -128 ? -128 : v > 127 ? 127 : v);
        }
        tlo[a] = _mm_loadu_si128((const __m128i*)lo);
        thi[a] = _mm_loadu_si128((const __m128i*)hi);
    }
    for (int64_t r = 0; r < m; ++r) B[r] = _mm256_set1_epi16(bias ? bias[r] : 0);
    const __m256i go_v = _mm256_set1_epi16((int16_t)go), ge_v = _mm256_set1_epi16((int16_t)ge);
    const __m256i zero = _mm256_setzero_si256(), neg = _mm256_set1_epi16(-16384);
    const __m128i fifteen = _mm_set1_epi8(15), low_nib = _mm_set1_epi8(0x0f);
    for (int64_t start = 0; start < nj; start += 16) {
        const int nl = (int)(nj - start < 16 ? nj - start : 16);
        int64_t max_len = 0;
        for (int l = 0; l < nl; ++l)
            if (jobs[start + l].n > max_len) max_len = jobs[start + l].n;
        for (int64_t r = 0; r < m; ++r) {
            H[r] = zero;
            E[r] = neg;
        }
        __m256i max_h = zero, profile[24];
        for (int64_t col = 0; col < max_len; ++col) {
            uint8_t resid[16];
            int16_t act[16];
            for (int l = 0; l < 16; ++l) {
                if (l < nl && col < jobs[start + l].n) {
                    resid[l] = (uint8_t)code(jobs[start + l].s[col]);
                    act[l] = -1;
                } else {
                    resid[l] = 22;
                    act[l] = 0;
                }
            }
            const __m256i active = _mm256_loadu_si256((const __m256i*)act);
            const __m128i rv = _mm_loadu_si128((const __m128i*)resid);
            const __m128i idx = _mm_and_si128(rv, low_nib), is_hi = _mm_cmpgt_epi8(rv, fifteen);
            for (int a = 0; a < 24; ++a)
                profile[a] = _mm256_cvtepi8_epi16(_mm_blendv_epi8(_mm_shuffle_epi8(tlo[a], idx),
                                                                  _mm_shuffle_epi8(thi[a], idx), is_hi));
            __m256i u_f = neg, u_h = zero, ul_h = zero;
            for (int64_t r = 0; r < m; ++r) {
                const __m256i p = _mm256_adds_epi16(profile[code(q[r])], B[r]);
                const __m256i ph = H[r];
                const __m256i e = _mm256_max_epi16(_mm256_subs_epi16(ph, go_v), _mm256_subs_epi16(E[r], ge_v));
                const __m256i f = _mm256_max_epi16(_mm256_subs_epi16(u_h, go_v), _mm256_subs_epi16(u_f, ge_v));
                const __m256i diag = _mm256_adds_epi16(ul_h, p);
                const __m256i h = _mm256_max_epi16(_mm256_max_epi16(e, f), _mm256_max_epi16(diag, zero));
                max_h = _mm256_max_epi16(max_h, _mm256_and_si256(h, active));
                u_f = f;
                ul_h = ph;
                u_h = h;
                H[r] = h;
                E[r] = e;
            }
        }
        int16_t res[16];
        _mm256_storeu_si256((__m256i*)res, max_h);
        for (int l = 0; l < nl; ++l) out[jobs[start + l].idx] = res[l];
    }
}

/* 8-bit pass; returns the number of subjects to re-score (their jobs are
 * written to redo), or -1 when the scheme does not fit 8-bit lanes */
static int64_t many_u8(const uint8_t* q, int64_t m, const int8_t* bias, const mjob* jobs, int64_t nj,
                       const int32_t* mat, int32_t go, int32_t ge, __m256i* H, __m256i* E, __m256i* B,
                       int32_t* out, mjob* redo) {
    int smin = 1 << 30, smax = -(1 << 30), bmin = 0, bmax = 0;
    for (int i = 0; i < 576; ++i) {
        if (mat[i] < smin) smin = mat[i];
        if (mat[i] > smax) smax = mat[i];
    }
    if (bias) {
        bmin = 127;
        bmax = -128;
        for (int64_t r = 0; r < m; ++r) {
            if (bias[r] < bmin) bmin = bias[r];
            if (bias[r] > bmax) bmax = bias[r];
        }
    }
    int t = smin + (bmin < 0 ? bmin : 0);
    const int off = -(t < 0 ? t : 0);
    const int pmax = smax + (bmax > 0 ? bmax : 0) + off;
    if (pmax >= 128 || off > 127 || go < 0 || go > 255 || ge < 0 || ge > 255) return -1;
    const int limit = 255 - pmax;
    __m256i tlo[24], thi[24];
    for (int a = 0; a < 24; ++a) {
        uint8_t t0[16], t1[16];
        memset(t1, 0, sizeof t1);
        for (int x = 0; x < 16; ++x) t0[x] = (uint8_t)(mat[a * 24 + x] + off);
        for (int x = 16; x < 24; ++x) t1[x - 16] = (uint8_t)(mat[a * 24 + x] + off);
        tlo[a] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)t0));
        thi[a] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)t1));
    }
    for (int64_t r = 0; r < m; ++r) B[r] = _mm256_set1_epi8(bias ? bias[r] : 0);
    const __m256i go_v = _mm256_set1_epi8((char)go), ge_v = _mm256_set1_epi8((char)ge);
    const __m256i offv = _mm256_set1_epi8((char)off), zero = _mm256_setzero_si256();
    const __m256i nib = _mm256_set1_epi8(0x0f), fifteen = _mm256_set1_epi8(15);
    int64_t nredo = 0;
    for (int64_t start = 0; start < nj; start += 32) {
        const int nl = (int)(nj - start < 32 ? nj - start : 32);
        int64_t max_len = 0;
        for (int l = 0; l < nl; ++l)
            if (jobs[start + l].n > max_len) max_len = jobs[start + l].n;
        for (int64_t r = 0; r < m; ++r) {
            H[r] = zero;
            E[r] = zero;
        }
        __m256i max_h = zero, profile[24];
        for (int64_t col = 0; col < max_len; ++col) {
            uint8_t resid[32], act[32];
            for (int l = 0; l < 32; ++l) {
                if (l < nl && col < jobs[start + l].n) {
                    resid[l] = (uint8_t)code(jobs[start + l].s[col]);
                    act[l] = 0xFF;
                } else {
                    resid[l] = 22;
                    act[l] = 0;
                }
            }
            const __m256i active = _mm256_loadu_si256((const __m256i*)act);
            const __m256i rv = _mm256_loadu_si256((const __m256i*)resid);
            const __m256i idx = _mm256_and_si256(rv, nib), ishi = _mm256_cmpgt_epi8(rv, fifteen);
            for (int a = 0; a < 24; ++a)
                profile[a] = _mm256_blendv_epi8(_mm256_shuffle_epi8(tlo[a], idx), _mm256_shuffle_epi8(thi[a], idx), ishi);
            __m256i u_f = zero, u_h = zero, ul_h = zero;
            for (int64_t r = 0; r < m; ++r) {
                const __m256i p = _mm256_add_epi8(profile[code(q[r])], B[r]);
                const __m256i ph = H[r];
                const __m256i e = _mm256_max_epu8(_mm256_subs_epu8(ph, go_v), _mm256_subs_epu8(E[r], ge_v));
                const __m256i f = _mm256_max_epu8(_mm256_subs_epu8(u_h, go_v), _mm256_subs_epu8(u_f, ge_v));
                const __m256i diag = _mm256_subs_epu8(_mm256_adds_epu8(ul_h, p), offv);
                const __m256i h = _mm256_max_epu8(_mm256_max_epu8(e, f), diag);
                max_h = _mm256_max_epu8(max_h, _mm256_and_si256(h, active));
                u_f = f;
                ul_h = ph;
                u_h = h;
                H[r] = h;
                E[r] = e;
            }
        }
        uint8_t res[32];
        _mm256_storeu_si256((__m256i*)res, max_h);
        for (int l = 0; l < nl; ++l) {
            if ((int)res[l] >= limit) redo[nredo++] = jobs[start + l];
            else out[jobs[start + l].idx] = res[l];
        }
    }
    return nredo;
}

typedef struct {
    const uint8_t *qres, *sres;
    const uint64_t *qoff, *soff;
    const int8_t* bias;
    uint32_t nq, ns;
    const int32_t* mat;
    int32_t go, ge;
    int32_t* out;          /* [nq][ns] */
    int64_t next;          /* next query (atomic) */
    mjob* order;           /* subjects longest first, shared */
} many_drv;

static void* many_worker(void* p) {
    many_drv* D = (many_drv*)p;
    int64_t maxm = 1;
    for (uint32_t i = 0; i < D->nq; ++i)
        if ((int64_t)(D->qoff[i + 1] - D->qoff[i]) > maxm) maxm = (int64_t)(D->qoff[i + 1] - D->qoff[i]);
    __m256i* H = (__m256i*)aligned_alloc(32, sizeof(__m256i) * maxm);
    __m256i* E = (__m256i*)aligned_alloc(32, sizeof(__m256i) * maxm);
    __m256i* B = (__m256i*)aligned_alloc(32, sizeof(__m256i) * maxm);
    mjob* redo = (mjob*)malloc(sizeof(mjob) * (D->ns ? D->ns : 1));
    int32_t* o16 = (int32_t*)malloc(sizeof(int32_t) * (D->ns ? D->ns : 1));
    for (;;) {
        const int64_t qi = __atomic_fetch_add(&D->next, 1, __ATOMIC_RELAXED);
        if (qi >= (int64_t)D->nq) break;
        const uint8_t* q = D->qres + D->qoff[qi];
        const int64_t m = (int64_t)(D->qoff[qi + 1] - D->qoff[qi]);
        const int8_t* b = D->bias ? D->bias + D->qoff[qi] : NULL;
        int32_t* out = D->out + (uint64_t)qi * D->ns;
        memset(out, 0, sizeof(int32_t) * D->ns);
        if (m == 0) continue;
        int64_t nr = many_u8(q, m, b, D->order, D->ns, D->mat, D->go, D->ge, H, E, B, out, redo);
        if (nr < 0) {
            many_i16(q, m, b, D->order, D->ns, D->mat, D->go, D->ge, H, E, B, out);
        } else if (nr > 0) {
            /* the redo list keeps the longest-first order */
            many_i16(q, m, b, redo, nr, D->mat, D->go, D->ge, H, E, B, out);
        }
        for (uint32_t s = 0; s < D->ns; ++s)
            if (out[s] >= 32767) {
                int32_t sc;
                int64_t qe, se;
                uint64_t cells;
                const int64_t n = (int64_t)(D->soff[s + 1] - D->soff[s]);
                or_banded_score(q, m, D->sres + D->soff[s], n, b, D->mat, D->go, D->ge, -m, n, &sc, &qe, &se, &cells);
                out[s] = sc;
            }
        (void)o16;
    }
    free(H);
    free(E);
    free(B);
    free(redo);
    free(o16);
    return NULL;
}

/* score_many of every query against every subject, queries spread over
 * nthreads threads; out [nq][ns]. */
void cpu_score_many_avx2(const uint8_t* qres, const uint64_t* qoff, uint32_t nq, const int8_t* bias,
                         const uint8_t* sres, const uint64_t* soff, uint32_t ns, const int32_t* mat, int32_t go,
                         int32_t ge, int32_t* out, int nthreads) {
    many_drv D;
    D.qres = qres;
    D.sres = sres;
    D.qoff = qoff;
    D.soff = soff;
    D.bias = bias;
    D.nq = nq;
    D.ns = ns;
    D.mat = mat;
    D.go = go;
    D.ge = ge;
    D.out = out;
    D.next = 0;
    D.order = (mjob*)malloc(sizeof(mjob) * (ns ? ns : 1));
    for (uint32_t s = 0; s < ns; ++s) {
        D.order[s].s = sres + soff[s];
        D.order[s].n = (int64_t)(soff[s + 1] - soff[s]);
        D.order[s].idx = s;
    }
    qsort(D.order, ns, sizeof(mjob), mjob_cmp);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, many_worker, &D);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(D.order);
}

/* ---------------- ungapped X-drop, 16 steps per vector ------------------- */
/*
 * scan.rs:934-1031 (xdrop_dir_avx2 / xdrop_avx2), restated: each direction
 * runs in blocks of 16 steps; a block's step scores (scalar table reads plus
 * the query bias) are summed and maximised with in-register prefix scans
 * relative to the score at the block start, and the first step whose sum
 * falls more than x below the running best ends the extension. Outputs equal
 * or_xdrop's (score, q_lo, q_hi, steps); hits whose scores could leave the
 * int16 lanes (|substitution score| > 1000 or x > 16000) take or_xdrop.
 */

/* 16-lane inclusive prefix sum / prefix max of int16 (in each 128-bit half,
 * then the low half's last lane carried into the high half). */
static inline __m256i carry_lo_half(__m256i v) {
    const __m256i b7 = _mm256_setr_epi8(14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14,
                                        15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15, 14, 15);
    return _mm256_shuffle_epi8(_mm256_permute2x128_si256(v, v, 0x08), b7);
}
static inline __m256i psum16(__m256i r) {
    r = _mm256_add_epi16(r, _mm256_slli_si256(r, 2));
    r = _mm256_add_epi16(r, _mm256_slli_si256(r, 4));
    r = _mm256_add_epi16(r, _mm256_slli_si256(r, 8));
    return _mm256_add_epi16(r, carry_lo_half(r));
}
static inline __m256i pmax16(__m256i m, __m256i floor) {
    m = _mm256_max_epi16(m, _mm256_slli_si256(m, 2));
    m = _mm256_max_epi16(m, _mm256_slli_si256(m, 4));
    m = _mm256_max_epi16(m, _mm256_slli_si256(m, 8));
    m = _mm256_max_epi16(m, carry_lo_half(m));
    return _mm256_max_epi16(m, floor);
}

/* One direction: n steps available; step t pairs q[d*t + o], s[d*t + o] with
 * (d, o) = (1, 0) rightwards and (-1, -1) leftwards. Returns the best sum,
 * the steps to it and the steps taken. */
static void xdrop_dir(const int16_t* T, const uint8_t* q, const int8_t* bias, const uint8_t* s, int64_t n, int left,
                      int32_t x, int32_t* best_out, int64_t* end_out, int64_t* taken_out) {
    int32_t cur = 0, best = 0;
    int64_t end = 0, t0 = 0;
    const __m256i xv = _mm256_set1_epi16((short)x);
    while (t0 < n) {
        int16_t sc[16] __attribute__((aligned(32)));
        const int64_t vk = n - t0 < 16 ? n - t0 : 16;
        for (int k = 0; k < 16; ++k) {
            if (k < vk) {
                const int64_t i = left ? -1 - (t0 + k) : t0 + k;
                sc[k] = (int16_t)(T[code(q[i]) * 32 + code(s[i])] + (bias ? bias[i] : 0));
            } else {
                sc[k] = 0;
            }
        }
        const __m256i r = psum16(_mm256_load_si256((const __m256i*)sc));
        const int32_t base = best - cur;
        const __m256i m = pmax16(r, _mm256_set1_epi16((short)base));
        const uint32_t valid = vk == 16 ? 0xFFFFFFFFu : ((1u << (2 * vk)) - 1u);
        const uint32_t stop = (uint32_t)_mm256_movemask_epi8(_mm256_cmpgt_epi16(_mm256_sub_epi16(m, xv), r)) & valid;
        const int last = stop ? (int)(__builtin_ctz(stop) / 2) : (int)vk - 1;
        int16_t ma[16], ra[16];
        _mm256_storeu_si256((__m256i*)ma, m);
        _mm256_storeu_si256((__m256i*)ra, r);
        const int32_t bt = ma[last];
        if (bt > base) {
            const uint32_t eq = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi16(r, _mm256_set1_epi16((short)bt)));
            end = t0 + __builtin_ctz(eq) / 2 + 1;
            best = cur + bt;
        }
        if (stop) {
            t0 += last + 1;
            break;
        }
        cur += ra[vk - 1];
        t0 += vk;
    }
    *best_out = best;
    *end_out = end;
    *taken_out = t0;
}

typedef struct {
    const uint8_t *qres, *sres;
    const uint64_t *qoff, *soff;
    const int8_t* bias;
    const uint32_t* hits;
    uint64_t nhits;
    const int32_t* mat;
    int16_t T[24 * 32];
    int wide;
    int32_t x;
    int32_t* score;
    uint32_t *lo, *hi;
    uint64_t* steps;
    int64_t next;
} xdrop_drv;

static void* xdrop_worker(void* p) {
    xdrop_drv* D = (xdrop_drv*)p;
    const int64_t chunk = 4096;
    for (;;) {
        const int64_t h0 = __atomic_fetch_add(&D->next, chunk, __ATOMIC_RELAXED);
        if (h0 >= (int64_t)D->nhits) break;
        const int64_t h1 = h0 + chunk < (int64_t)D->nhits ? h0 + chunk : (int64_t)D->nhits;
        for (int64_t h = h0; h < h1; ++h) {
            const uint32_t* H = D->hits + 4 * h;
            const uint8_t* q = D->qres + D->qoff[H[0]];
            const int64_t m = (int64_t)(D->qoff[H[0] + 1] - D->qoff[H[0]]);
            const uint8_t* s = D->sres + D->soff[H[1]];
            const int64_t n = (int64_t)(D->soff[H[1] + 1] - D->soff[H[1]]);
            const int8_t* b = D->bias ? D->bias + D->qoff[H[0]] : NULL;
            const int64_t qp = H[2], sp = H[3];
            int32_t sc;
            int64_t lo, hi;
            uint64_t st;
            if (D->wide) {
                or_xdrop(q, m, b, s, n, D->mat, qp, sp, D->x, &sc, &lo, &hi, &st);
            } else {
                const int64_t nr = (m - qp) < (n - sp) ? (m - qp) : (n - sp);
                const int64_t nl = qp < sp ? qp : sp;
                int32_t br, bl;
                int64_t er, el, tr, tl;
                xdrop_dir(D->T, q + qp, b ? b + qp : NULL, s + sp, nr, 0, D->x, &br, &er, &tr);
                xdrop_dir(D->T, q + qp, b ? b + qp : NULL, s + sp, nl, 1, D->x, &bl, &el, &tl);
                sc = br + bl;
                lo = qp - el;
                hi = qp + er;
                st = (uint64_t)(tr + tl);
            }
            D->score[h] = sc;
            D->lo[h] = (uint32_t)lo;
            D->hi[h] = (uint32_t)hi;
            D->steps[h] = st;
        }
    }
    return NULL;
}

/* X-drop of every hit (query, subject, q_pos, s_pos) over nthreads threads. */
void cpu_xdrop_avx2(const uint8_t* qres, const uint64_t* qoff, const int8_t* bias, const uint8_t* sres,
                    const uint64_t* soff, const uint32_t* hits, uint64_t nhits, const int32_t* mat, int32_t x,
                    int32_t* score, uint32_t* lo, uint32_t* hi, uint64_t* steps, int nthreads) {
    xdrop_drv* D = (xdrop_drv*)calloc(1, sizeof(xdrop_drv));
    D->qres = qres;
    D->qoff = qoff;
    D->bias = bias;
    D->sres = sres;
    D->soff = soff;
    D->hits = hits;
    D->nhits = nhits;
    D->mat = mat;
    D->x = x;
    D->score = score;
    D->lo = lo;
    D->hi = hi;
    D->steps = steps;
    D->wide = x > 16000 || x < 0;
    for (int a = 0; a < 24; ++a)
        for (int c = 0; c < 32; ++c) {
            const int32_t v = mat[a * 24 + (c < 24 ? c : 22)];
            if (v < -1000 || v > 1000) D->wide = 1;
            D->T[a * 32 + c] = (int16_t)v;
        }
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, xdrop_worker, D);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(D);
}
