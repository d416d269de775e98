/*
 * CPU oracle for the candidate-scoring hot path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the contract, the reference lines each function restates
 * and how each one is pinned. Nothing in the product links this file.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define AA_X 22

int or_code_index(uint8_t c) { return c < 24 ? (int)c : AA_X; }

static inline int32_t S(const int32_t* mat, uint8_t a, uint8_t b) {
    return mat[or_code_index(a) * 24 + or_code_index(b)];
}

/* reference.rs:15-52 */
void or_sw_reference(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                     const int32_t* mat, const int8_t* bias, int64_t go, int64_t ge,
                     int64_t* score, int64_t* q_end, int64_t* s_end) {
    *score = 0;
    *q_end = 0;
    *s_end = 0;
    if (m == 0 || n == 0) return;
    const int64_t neg = INT64_MIN / 4;
    const int64_t w = n + 1;
    int64_t* h = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    int64_t* e = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    for (int64_t t = 0; t < (m + 1) * w; t++) {
        h[t] = 0;
        e[t] = neg;
        f[t] = neg;
    }
    int64_t best = 0, bi = 0, bj = 0;
    for (int64_t i = 1; i <= m; i++) {
        int64_t b = bias ? bias[i - 1] : 0;
        for (int64_t j = 1; j <= n; j++) {
            int64_t ee = h[i * w + j - 1] - go;
            int64_t t2 = e[i * w + j - 1] - ge;
            if (t2 > ee) ee = t2;
            int64_t ff = h[(i - 1) * w + j] - go;
            t2 = f[(i - 1) * w + j] - ge;
            if (t2 > ff) ff = t2;
            int64_t dd = h[(i - 1) * w + j - 1] + S(mat, q[i - 1], s[j - 1]) + b;
            int64_t v = 0;
            if (ee > v) v = ee;
            if (ff > v) v = ff;
            if (dd > v) v = dd;
            e[i * w + j] = ee;
            f[i * w + j] = ff;
            h[i * w + j] = v;
            if (v > best) {
                best = v;
                bi = i;
                bj = j;
            }
        }
    }
    free(h);
    free(e);
    free(f);
    *score = best;
    *q_end = bi;
    *s_end = bj;
}

#define NEG32 (INT32_MIN / 4)

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* banded.rs:123-131: band cell k of row i is column j = i + lo + k. */
static int row_range(int64_t i, int64_t lo, int64_t w, int64_t n, int64_t* kmin, int64_t* kmax) {
    int64_t a = -(i + lo);
    if (a < 0) a = 0;
    int64_t b = n - 1 - i - lo;
    if (b > w - 1) b = w - 1;
    *kmin = a;
    *kmax = b;
    return a <= b;
}

/* banded.rs:134-224 (scalar kernel), result convention of banded.rs:408-449. */
void or_banded_score(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                     const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge,
                     int64_t band_lo, int64_t band_hi,
                     int32_t* score, int64_t* q_end, int64_t* s_end, uint64_t* cells) {
    *score = 0;
    *q_end = 0;
    *s_end = 0;
    *cells = 0;
    if (m == 0 || n == 0 || band_hi < band_lo) return;
    /* Band::clip (banded.rs:45-47) */
    int64_t lo = band_lo > -(m - 1) ? band_lo : -(m - 1);
    int64_t hi = band_hi < n - 1 ? band_hi : n - 1;
    int64_t w = hi - lo + 1;
    if (w <= 0) return;
    int32_t* hp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* fp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* hc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* fc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    for (int64_t k = 0; k <= w; k++) hp[k] = fp[k] = hc[k] = fc[k] = NEG32;
    int32_t best = 0;
    int64_t bq = 0, bs = 0;
    uint64_t nc = 0;
    for (int64_t i = 0; i < m; i++) {
        int64_t kmin, kmax;
        if (!row_range(i, lo, w, n, &kmin, &kmax)) {
            for (int64_t k = 0; k < w; k++) hc[k] = fc[k] = NEG32;
        } else {
            int32_t b = bias ? bias[i] : 0;
            int32_t e = NEG32, hl = NEG32;
            for (int64_t k = kmin; k <= kmax; k++) {
                int64_t j = i + lo + k;
                int32_t ecand = imax(hl - go, e - ge);
                e = k > kmin ? ecand : NEG32;
                int32_t f = i >= 1 ? imax(hp[k + 1] - go, fp[k + 1] - ge) : NEG32;
                int32_t hd = (i >= 1 && j >= 1) ? hp[k] : 0;
                int32_t diag = hd + S(mat, q[i], s[j]) + b;
                int32_t h = imax(imax(diag, e), imax(f, 0));
                hc[k] = h;
                fc[k] = f;
                hl = h;
                if (h > best) {
                    best = h;
                    bq = i + 1;
                    bs = j + 1;
                }
            }
            nc += (uint64_t)(kmax - kmin + 1);
            for (int64_t k = 0; k < kmin; k++) hc[k] = fc[k] = NEG32;
            for (int64_t k = kmax + 1; k < w; k++) hc[k] = fc[k] = NEG32;
        }
        int32_t* t = hp;
        hp = hc;
        hc = t;
        t = fp;
        fp = fc;
        fc = t;
    }
    free(hp);
    free(fp);
    free(hc);
    free(fc);
    *cells = nc;
    if (best > 0) {
        *score = best;
        *q_end = bq;
        *s_end = bs;
    }
}

/* banded.rs:227-333 with the AVX2 row stop (banded_simd.rs:324, 594), walk
 * banded.rs:336-403. */
int or_band_align(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n, const int8_t* bias,
                  const int32_t* mat, int32_t go, int32_t ge, int64_t band_lo, int64_t band_hi,
                  int32_t stop_at, uint64_t max_dir_bytes, int32_t* score, int64_t* q_start,
                  int64_t* q_end, int64_t* s_start, int64_t* s_end, uint64_t* cells,
                  uint8_t* op_codes, uint32_t* op_lens, int64_t* nops) {
    *score = 0;
    *q_start = *q_end = *s_start = *s_end = 0;
    *cells = 0;
    *nops = 0;
    if (m == 0 || n == 0 || band_hi < band_lo) return 1;
    int64_t lo = band_lo > -(m - 1) ? band_lo : -(m - 1);
    int64_t hi = band_hi < n - 1 ? band_hi : n - 1;
    int64_t w = hi - lo + 1;
    if (w <= 0) return 1;
    if ((uint64_t)m * (uint64_t)w > max_dir_bytes) return 2;
    uint8_t* dir = (uint8_t*)calloc((size_t)(m * w), 1);
    int32_t* hp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* fp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* hc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    int32_t* fc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w + 1));
    for (int64_t k = 0; k <= w; k++) hp[k] = fp[k] = hc[k] = fc[k] = NEG32;
    int32_t best = 0;
    int64_t bi = 0, bk = 0;
    uint64_t nc = 0;
    for (int64_t i = 0; i < m; i++) {
        int64_t kmin, kmax;
        const int has_cells = row_range(i, lo, w, n, &kmin, &kmax);
        if (!has_cells) {
            for (int64_t k = 0; k < w; k++) hc[k] = fc[k] = NEG32;
        } else {
            const int32_t b = bias ? bias[i] : 0;
            int32_t e = NEG32, hl = NEG32;
            uint8_t* drow = dir + i * w;
            for (int64_t k = kmin; k <= kmax; k++) {
                const int64_t j = i + lo + k;
                const int32_t eo = hl - go, ee = e - ge;
                const int e_ext = ee > eo;
                e = k > kmin ? imax(ee, eo) : NEG32;
                const int32_t fo = hp[k + 1] - go, fe = fp[k + 1] - ge;
                const int f_ext = fe > fo;
                const int32_t f = i >= 1 ? imax(fe, fo) : NEG32;
                const int32_t hd = (i >= 1 && j >= 1) ? hp[k] : 0;
                const int32_t diag = hd + S(mat, q[i], s[j]) + b;
                const int c1 = diag > 0;
                const int32_t h1 = imax(diag, 0);
                const int c2 = e > h1;
                const int32_t h2 = imax(h1, e);
                const int c3 = f > h2;
                const int32_t h = imax(h2, f);
                const int org = c3 ? 3 : (c2 ? 2 : c1);
                hc[k] = h;
                fc[k] = f;
                drow[k] = (uint8_t)(org | (e_ext << 2) | (f_ext << 3));
                hl = h;
                if (h > best) {
                    best = h;
                    bi = i;
                    bk = k;
                }
            }
            nc += (uint64_t)(kmax - kmin + 1);
            for (int64_t k = 0; k < kmin; k++) hc[k] = fc[k] = NEG32;
            for (int64_t k = kmax + 1; k < w; k++) hc[k] = fc[k] = NEG32;
        }
        int32_t* t = hp;
        hp = hc;
        hc = t;
        t = fp;
        fp = fc;
        fc = t;
        /* the AVX2 kernels test the stop after rows holding band cells only */
        if (has_cells && best >= stop_at) break;
    }
    free(hp);
    free(fp);
    free(hc);
    free(fc);
    *cells = nc;
    if (best <= 0) {
        free(dir);
        return 1;
    }
    /* walk: ops collected from the end cell backwards */
    uint8_t* rev = (uint8_t*)malloc((size_t)(m + n + 1));
    int64_t nrev = 0;
    int64_t i = bi, k = bk;
    int state = 0; /* 0 = H, 1 = E (query gap, 'D'), 2 = F (subject gap, 'I') */
    for (;;) {
        const uint8_t d = dir[i * w + k];
        if (state == 0) {
            const int org = d & 3;
            if (org == 0) break;
            if (org == 1) {
                rev[nrev++] = 'M';
                const int64_t j = i + lo + k;
                if (i == 0 || j == 0) break;
                i -= 1;
            } else if (org == 2) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            rev[nrev++] = 'D';
            k -= 1;
            if (!(d & 4)) state = 0;
        } else {
            rev[nrev++] = 'I';
            i -= 1;
            k += 1;
            if (!(d & 8)) state = 0;
        }
    }
    const int64_t qe = bi + 1, se = bi + lo + bk + 1;
    int64_t qs = qe, ss = se, nr = 0;
    for (int64_t x = nrev - 1; x >= 0; x--) {
        const uint8_t op = rev[x];
        if (op == 'M') {
            qs--;
            ss--;
        } else if (op == 'I') {
            qs--;
        } else {
            ss--;
        }
        if (nr > 0 && op_codes[nr - 1] == op) {
            op_lens[nr - 1]++;
        } else {
            op_codes[nr] = op;
            op_lens[nr] = 1;
            nr++;
        }
    }
    free(rev);
    free(dir);
    *score = best;
    *q_start = qs;
    *q_end = qe;
    *s_start = ss;
    *s_end = se;
    *nops = nr;
    return 0;
}

/* Independent (i, j)-coordinate formulation: H/E/F over the whole (m+1)x(n+1)
 * matrix, cells outside the band (or its clip) are unreachable. */
void or_banded_score_ij(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                        const int8_t* bias, const int32_t* mat, int32_t go, int32_t ge,
                        int64_t band_lo, int64_t band_hi,
                        int32_t* score, int64_t* q_end, int64_t* s_end) {
    *score = 0;
    *q_end = 0;
    *s_end = 0;
    if (m == 0 || n == 0 || band_hi < band_lo) return;
    const int64_t neg = INT64_MIN / 4;
    const int64_t w = n + 1;
    int64_t* h = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    int64_t* e = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * (size_t)((m + 1) * w));
    /* Row 0 / column 0 of the (m+1)x(n+1) matrix are the border: H = 0 there
     * (a local alignment may start anywhere), E = F = -inf. */
    for (int64_t t = 0; t < (m + 1) * w; t++) {
        h[t] = 0;
        e[t] = neg;
        f[t] = neg;
    }
    int64_t best = 0, bi = 0, bj = 0;
    for (int64_t i = 1; i <= m; i++) {
        for (int64_t j = 1; j <= n; j++) {
            int64_t d = (j - 1) - (i - 1);
            if (d < band_lo || d > band_hi) {
                h[i * w + j] = neg;
                e[i * w + j] = neg;
                f[i * w + j] = neg;
                continue;
            }
            int64_t ee = h[i * w + j - 1] - go, t2 = e[i * w + j - 1] - ge;
            if (t2 > ee) ee = t2;
            int64_t ff = h[(i - 1) * w + j] - go;
            t2 = f[(i - 1) * w + j] - ge;
            if (t2 > ff) ff = t2;
            int64_t hd = h[(i - 1) * w + j - 1];
            int64_t dd = hd + S(mat, q[i - 1], s[j - 1]) + (bias ? bias[i - 1] : 0);
            int64_t v = 0;
            if (dd > v) v = dd;
            if (ee > v) v = ee;
            if (ff > v) v = ff;
            h[i * w + j] = v;
            e[i * w + j] = ee;
            f[i * w + j] = ff;
            if (v > best) {
                best = v;
                bi = i;
                bj = j;
            }
        }
    }
    free(h);
    free(e);
    free(f);
    *score = (int32_t)best;
    *q_end = bi;
    *s_end = bj;
}

struct many_args {
    const uint8_t *qres, *sres;
    const int8_t* bias;
    const uint64_t* qoff;
    const uint32_t* qlen;
    const uint64_t* soff;
    const uint32_t* slen;
    const int32_t *lo, *hi;
    int64_t begin, end;
    const int32_t* mat;
    int32_t go, ge;
    int32_t* score;
    uint32_t *q_end, *s_end;
    uint64_t* cells;
};

static void* many_worker(void* p) {
    struct many_args* a = (struct many_args*)p;
    for (int64_t k = a->begin; k < a->end; k++) {
        int32_t sc;
        int64_t qe, se;
        uint64_t c;
        or_banded_score(a->qres + a->qoff[k], a->qlen[k], a->sres + a->soff[k], a->slen[k],
                        a->bias ? a->bias + a->qoff[k] : NULL, a->mat, a->go, a->ge, a->lo[k],
                        a->hi[k], &sc, &qe, &se, &c);
        a->score[k] = sc;
        if (a->q_end) a->q_end[k] = (uint32_t)qe;
        if (a->s_end) a->s_end[k] = (uint32_t)se;
        if (a->cells) a->cells[k] = c;
    }
    return NULL;
}

void or_banded_many(const uint8_t* qres, const uint8_t* sres, const int8_t* bias,
                    const uint64_t* qoff, const uint32_t* qlen,
                    const uint64_t* soff, const uint32_t* slen,
                    const int32_t* lo, const int32_t* hi, int64_t njobs,
                    const int32_t* mat, int32_t go, int32_t ge,
                    int32_t* score, uint32_t* q_end, uint32_t* s_end, uint64_t* cells,
                    int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    struct many_args args[256];
    for (int t = 0; t < nthreads; t++) {
        struct many_args a = {qres, sres, bias, qoff, qlen, soff, slen, lo, hi,
                              njobs * t / nthreads, njobs * (t + 1) / nthreads,
                              mat, go, ge, score, q_end, s_end, cells};
        args[t] = a;
        pthread_create(&th[t], NULL, many_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* scan.rs:859-904 (xdrop_cat): right and left extensions, each stopping when
 * the running sum falls more than x below its best; best and its step are
 * updated on strict improvement only. */
void or_xdrop(const uint8_t* q, int64_t m, const int8_t* bias, const uint8_t* s, int64_t n,
              const int32_t* mat, int64_t qp, int64_t sp, int32_t x,
              int32_t* score, int64_t* q_lo, int64_t* q_hi, uint64_t* steps) {
    int64_t nr = (m - qp) < (n - sp) ? (m - qp) : (n - sp);
    int64_t nl = qp < sp ? qp : sp;
    int32_t best_r = 0, cur = 0;
    int64_t end_t = 0, tr = 0;
    while (tr < nr) {
        cur += S(mat, q[qp + tr], s[sp + tr]) + (bias ? bias[qp + tr] : 0);
        tr++;
        if (cur > best_r) {
            best_r = cur;
            end_t = tr;
        }
        if (cur < best_r - x) break;
    }
    int32_t best_l = 0;
    cur = 0;
    int64_t st_t = 0, tl = 0;
    while (tl < nl) {
        tl++;
        cur += S(mat, q[qp - tl], s[sp - tl]) + (bias ? bias[qp - tl] : 0);
        if (cur > best_l) {
            best_l = cur;
            st_t = tl;
        }
        if (cur < best_l - x) break;
    }
    *score = best_r + best_l;
    *q_lo = qp - st_t;
    *q_hi = qp + end_t;
    if (steps) *steps = (uint64_t)(tr + tl);
}

/* engine.rs:2460-2494 */
void or_xdrop_scalar_twin(const uint8_t* q, int64_t m, const int8_t* bias, const uint8_t* s,
                          int64_t n, const int32_t* mat, int64_t qp, int64_t sp, int32_t x,
                          int32_t* score, int64_t* q_lo, int64_t* q_hi) {
    int32_t best_r = 0, cur = 0;
    int64_t end = qp, i = qp, j = sp;
    while (i < m && j < n) {
        cur += S(mat, q[i], s[j]) + (bias ? bias[i] : 0);
        i++;
        j++;
        if (cur > best_r) {
            best_r = cur;
            end = i;
        } else if (cur < best_r - x) {
            break;
        }
    }
    int32_t best_l = 0;
    cur = 0;
    int64_t start = qp;
    i = qp - 1;
    j = sp - 1;
    while (i >= 0 && j >= 0) {
        cur += S(mat, q[i], s[j]) + (bias ? bias[i] : 0);
        if (cur > best_l) {
            best_l = cur;
            start = i;
        } else if (cur < best_l - x) {
            break;
        }
        i--;
        j--;
    }
    *score = best_r + best_l;
    *q_lo = start;
    *q_hi = end > qp ? end : qp;
}

/* reference.rs:106-119 */
int64_t or_best_ungapped_on_diagonal(const uint8_t* q, int64_t m, const uint8_t* s, int64_t n,
                                     int64_t d, const int32_t* mat, const int8_t* bias) {
    int64_t best = 0, cur = 0;
    for (int64_t i = 0; i < m; i++) {
        int64_t j = i + d;
        if (j < 0 || j >= n) continue;
        cur += S(mat, q[i], s[j]) + (bias ? bias[i] : 0);
        if (cur < 0) cur = 0;
        if (cur > best) best = cur;
    }
    return best;
}

/* bias.rs:32-53 Robinson & Robinson background frequencies per 1000. */
static const double ROBINSON_PER_1000[20] = {78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95,
                                             73.77, 21.99, 51.42, 90.19, 57.44, 22.43, 38.56,
                                             52.03, 71.20, 58.41, 13.30, 32.16, 64.41};

/* bias.rs:56-122. Sums run left to right as in the reference so the f64
 * values (and hence the rounding) are identical. */
void or_composition_bias(const uint8_t* q, int64_t n, const int32_t* mat, int8_t* out) {
    for (int64_t i = 0; i < n; i++) out[i] = 0;
    if (n < 2) return;
    double tot = 0.0;
    for (int a = 0; a < 20; a++) tot += ROBINSON_PER_1000[a];
    double p[20], bg[20];
    for (int a = 0; a < 20; a++) p[a] = ROBINSON_PER_1000[a] / tot;
    for (int r = 0; r < 20; r++) {
        double acc = 0.0;
        for (int a = 0; a < 20; a++) acc += p[a] * (double)mat[r * 24 + a];
        bg[r] = acc;
    }
    int32_t acc[20];
    memset(acc, 0, sizeof(acc));
    const int64_t half = 20;
    int64_t lo = 0, hi = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t want_hi = i + half + 1 < n ? i + half + 1 : n;
        int64_t want_lo = i > half ? i - half : 0;
        while (hi < want_hi) {
            int c = q[hi] < 23 ? q[hi] : 23;
            for (int r = 0; r < 20; r++) acc[r] += mat[r * 24 + c];
            hi++;
        }
        while (lo < want_lo) {
            int c = q[lo] < 23 ? q[lo] : 23;
            for (int r = 0; r < 20; r++) acc[r] -= mat[r * 24 + c];
            lo++;
        }
        int r = q[i];
        if (r < 20) {
            int32_t w = (int32_t)(hi - lo);
            if (w >= 2) {
                int32_t own = mat[r * 24 + r];
                double neigh = (double)(acc[r] - own) / (double)(w - 1);
                double f = bg[r] - neigh;
                int32_t v = (int32_t)(f < 0.0 ? f - 0.5 : f + 0.5);
                if (v < -127) v = -127;
                if (v > 127) v = 127;
                out[i] = (int8_t)v;
            }
        }
    }
}
