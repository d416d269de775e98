/*
 * CPU oracle of the scan schedule -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates, scalar and single-threaded, what Engine::search_scan does from
 * the query batch to each query's candidate regions (SABER 1.4.0):
 *   build_batch            src/search/engine/scan.rs:1500-1596 (seed tables:
 *                          every query window of every shape whose care
 *                          residues are seedable, over-frequent codes skipped)
 *   over-frequent codes    scan.rs:1125-1230 with auto_occ_cap
 *                          (src/search/index.rs:319-336)
 *   scan_hits + filter     scan.rs:1603-1804 and filter_scalar :1812-1856
 *                          (32-residue window identity filter: seed-letter
 *                          identities + kx x exact identities >= min_ids)
 *   visit_chunk_inner      scan.rs:2084-2225 (survivors of a subject ordered by
 *                          batch position, stable, i.e. (query position,
 *                          shape, subject position); one group per query)
 *   visit_scan             scan.rs:2384-2431 (diagonal coverage, X-drop)
 *   finish_subject_into    scan.rs:2437-2474 (join, region threshold, top
 *                          regions per subject)
 *   per-query cap          scan.rs:1453-1466 (region_order, engine.rs:493-495)
 * Masking is off (no query or database seed mask). The table layouts of the
 * reference (dense / blocked / sketch) are lookup structures only; here a
 * sorted (code, batch position) list with binary search gives the same hits
 * in the same order.
 *
 * Pinned by tests/test_scan_oracle.py: brute-force hit enumeration over care
 * positions on tiny inputs (every (shape, query position, subject position)
 * whose reduced letters agree at every care position), the filter value
 * evaluated directly from its definition, the frequent-code rule from direct
 * counts, the visit replayed from or_xdrop results, and the region rules
 * against hand-worked cases.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PAD 32          /* scan.rs:37 */
#define NONSEED 0x7F    /* scan.rs:39 */
#define QSENT 0xFE      /* scan.rs:44 */
#define OCC_CAP_FLOOR 256u  /* index.rs:321 */

typedef struct {
    uint32_t code, g;
} Entry;

typedef struct {
    uint32_t g, t, shape;  /* batch position, subject position, shape */
} Surv;

typedef struct {
    int64_t diag;
    uint32_t q_lo, q_hi;
    int32_t score;
} Hsp;

typedef struct {
    or_region* v;
    uint64_t n, cap;
} RegionList;

typedef struct {
    uint32_t* q;   /* query, subject, qp, sp, shape per survivor */
    uint64_t n, cap;
} SurvLog;

typedef struct {
    RegionList* per_query;
    uint32_t nq;
    SurvLog log;
} ScanResult;

static void* xrealloc(void* p, size_t n) {
    void* r = realloc(p, n ? n : 1);
    if (!r) abort();
    return r;
}

static int cmp_entry(const void* a, const void* b) {
    const Entry *x = (const Entry*)a, *y = (const Entry*)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->g < y->g ? -1 : (x->g > y->g);
}

static int cmp_hsp_diag(const void* a, const void* b) {
    const Hsp *x = (const Hsp*)a, *y = (const Hsp*)b;
    return x->diag < y->diag ? -1 : (x->diag > y->diag);
}

/* region_order (engine.rs:493-495): best desc, subject, d_min, d_max */
static int cmp_region_order(const void* a, const void* b) {
    const or_region *x = (const or_region*)a, *y = (const or_region*)b;
    if (x->best != y->best) return x->best > y->best ? -1 : 1;
    if (x->subject != y->subject) return x->subject < y->subject ? -1 : 1;
    if (x->d_min != y->d_min) return x->d_min < y->d_min ? -1 : 1;
    return x->d_max < y->d_max ? -1 : (x->d_max > y->d_max);
}

/* finish_subject_into's order: best desc, d_min asc */
static int cmp_best_dmin(const void* a, const void* b) {
    const or_region *x = (const or_region*)a, *y = (const or_region*)b;
    if (x->best != y->best) return x->best > y->best ? -1 : 1;
    return x->d_min < y->d_min ? -1 : (x->d_min > y->d_min);
}

/* seed_code (scan.rs:811-823): None when a care residue is not a seed letter */
static int seed_code(const uint8_t* sb, const uint16_t* offs, uint32_t w, uint32_t asize, uint32_t* code) {
    uint32_t c = 0;
    for (uint32_t k = 0; k < w; ++k) {
        const uint8_t r = sb[offs[k]];
        if (r >= asize) return 0;
        c = c * asize + r;
    }
    *code = c;
    return 1;
}

/* auto_occ_cap (src/search/index.rs:329-336) */
uint64_t or_auto_occ_cap(uint64_t total, uint64_t ncodes, uint64_t db_residues, uint64_t ref_residues) {
    uint64_t floor = OCC_CAP_FLOOR;
    if (ref_residues > 0 && db_residues > ref_residues) {
        double f = (double)OCC_CAP_FLOOR * (double)db_residues / (double)ref_residues;
        floor = (uint64_t)f;
        if ((double)floor < f) floor += 1;
    }
    double v = (double)total / (double)(ncodes ? ncodes : 1) * 64.0;
    uint64_t c = (uint64_t)v;
    if ((double)c < v) c += 1;
    return c > floor ? c : floor;
}

static void surv_log(SurvLog* L, uint32_t q, uint32_t s, uint32_t qp, uint32_t sp, uint32_t shape) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->q = (uint32_t*)xrealloc(L->q, sizeof(uint32_t) * 5 * L->cap);
    }
    uint32_t* e = L->q + 5 * L->n++;
    e[0] = q, e[1] = s, e[2] = qp, e[3] = sp, e[4] = shape;
}

static void region_push(RegionList* R, const or_region* r) {
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 64;
        R->v = (or_region*)xrealloc(R->v, sizeof(or_region) * R->cap);
    }
    R->v[R->n++] = *r;
}

void* or_scan(const uint8_t* qres, const uint64_t* qoff, uint32_t nq, const int8_t* qbias,
              const uint8_t* sres, const uint64_t* soff, uint32_t ns, const int32_t* mat,
              const int32_t* region_thr, const or_scan_params* P, int log_survivors,
              or_scan_counters* ctr) {
    memset(ctr, 0, sizeof(*ctr));
    ScanResult* res = (ScanResult*)calloc(1, sizeof(ScanResult));
    res->nq = nq;
    res->per_query = (RegionList*)calloc(nq ? nq : 1, sizeof(RegionList));
    const uint32_t asize = P->asize;
    /* seed bytes of an alignment code: reduced letter or NONSEED (scan.rs:1072-1076) */
    uint8_t seed_map[256];
    for (int c = 0; c < 256; ++c) seed_map[c] = NONSEED;
    for (int c = 0; c < 24; ++c) seed_map[c] = P->alpha_map[c] == 255 ? NONSEED : P->alpha_map[c];

    /* batch layout (build_batch): query q at base[q], PAD bytes between */
    uint64_t* base = (uint64_t*)malloc(sizeof(uint64_t) * (nq + 1));
    uint64_t cur = PAD;
    for (uint32_t q = 0; q < nq; ++q) {
        base[q] = cur;
        cur += (qoff[q + 1] - qoff[q]) + PAD;
    }
    const uint64_t total = cur;
    uint8_t* red = (uint8_t*)malloc(total);    /* seed letter or QSENT */
    uint8_t* win = (uint8_t*)malloc(total);    /* alignment code of a seedable residue, else QSENT */
    memset(red, QSENT, total);
    memset(win, QSENT, total);
    uint64_t nres_q = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        const uint64_t m = qoff[q + 1] - qoff[q];
        nres_q += m;
        for (uint64_t i = 0; i < m; ++i) {
            const uint8_t c = qres[qoff[q] + i];
            const uint8_t r = P->alpha_map[c < 23 ? c : 23];
            if (r != 255) {
                red[base[q] + i] = r;
                win[base[q] + i] = c;
            }
        }
    }
    /* subject seed bytes */
    const uint64_t nres = ns ? soff[ns] : 0;
    uint8_t* sseed = (uint8_t*)malloc(nres + 1);
    for (uint64_t i = 0; i < nres; ++i) sseed[i] = seed_map[sres[i]];

    /* seed tables, one per shape */
    Entry** tab = (Entry**)calloc(P->nshapes, sizeof(Entry*));
    uint64_t* ntab = (uint64_t*)calloc(P->nshapes, sizeof(uint64_t));
    uint32_t cmin_span = 0xFFFFFFFFu;
    for (uint32_t si = 0; si < P->nshapes; ++si) {
        const uint16_t* offs = P->offsets + 16 * si;
        const uint32_t w = P->weight[si], span = P->span[si];
        if (span < cmin_span) cmin_span = span;
        uint64_t ncodes = 1;
        for (uint32_t k = 0; k < w; ++k) ncodes *= asize;
        /* database counts of every code (no mask), the frequent-code rule */
        uint32_t* cnt = (uint32_t*)calloc(ncodes, sizeof(uint32_t));
        uint64_t tot = 0;
        for (uint32_t s = 0; s < ns; ++s) {
            const uint64_t sl = soff[s + 1] - soff[s];
            if (sl < span) continue;
            for (uint64_t p = 0; p + span <= sl; ++p) {
                uint32_t c;
                if (seed_code(sseed + soff[s] + p, offs, w, asize, &c)) {
                    cnt[c]++;
                    tot++;
                }
            }
        }
        const uint64_t cap = P->seed_max_occ > 0 ? P->seed_max_occ
                                                 : or_auto_occ_cap(tot, ncodes, nres, P->ref_db_residues);
        Entry* e = NULL;
        uint64_t n = 0, ecap = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            const uint64_t m = qoff[q + 1] - qoff[q];
            if (m < span) continue;
            for (uint64_t p = 0; p + span <= m; ++p) {
                uint32_t c;
                if (!seed_code(red + base[q] + p, offs, w, asize, &c)) continue;
                if (tot > cap && cnt[c] > cap) {
                    ctr->seeds_frequent_skipped++;
                    continue;
                }
                if (n == ecap) {
                    ecap = ecap ? 2 * ecap : 1024;
                    e = (Entry*)xrealloc(e, sizeof(Entry) * ecap);
                }
                e[n].code = c;
                e[n].g = (uint32_t)(base[q] + p);
                ++n;
            }
        }
        free(cnt);
        qsort(e, n, sizeof(Entry), cmp_entry);
        tab[si] = e;
        ntab[si] = n;
        ctr->seed_positions += n;
    }
    const int64_t c_half = (int64_t)(cmin_span / 2);

    /* query of a batch position */
    uint32_t* qof = (uint32_t*)malloc(sizeof(uint32_t) * total);
    for (uint64_t g = 0; g < total; ++g) qof[g] = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < nq; ++q)
        for (uint64_t i = 0; i < qoff[q + 1] - qoff[q]; ++i) qof[base[q] + i] = q;

    Surv* sv = NULL;
    uint64_t nsv = 0, svcap = 0;
    Hsp* hsps = NULL;
    uint64_t hcap = 0;
    or_region* regs = NULL;
    uint64_t rcap = 0;
    uint32_t* dcov = NULL;
    uint64_t dcap = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        const uint64_t sb = soff[s], sl = soff[s + 1] - soff[s];
        nsv = 0;
        for (uint32_t si = 0; si < P->nshapes; ++si) {
            const uint16_t* offs = P->offsets + 16 * si;
            const uint32_t w = P->weight[si], span = P->span[si];
            if (ntab[si] == 0 || sl < span) continue;
            for (uint64_t p = 0; p + span <= sl; ++p) {
                uint32_t c;
                if (!seed_code(sseed + sb + p, offs, w, asize, &c)) continue;
                /* entries of this code: lower bound by binary search */
                uint64_t lo = 0, hi = ntab[si];
                while (lo < hi) {
                    const uint64_t mid = (lo + hi) / 2;
                    if (tab[si][mid].code < c) lo = mid + 1;
                    else hi = mid;
                }
                if (lo >= ntab[si] || tab[si][lo].code != c) continue;
                ctr->scan_candidates++;
                for (uint64_t k = lo; k < ntab[si] && tab[si][k].code == c; ++k) {
                    const uint32_t g = tab[si][k].g;
                    ctr->hits++;
                    if (P->min_ids > 0) {
                        /* filter_scalar: window of 32 starting at p + c - 16 */
                        const int64_t ss = (int64_t)p + c_half - 16;
                        const int64_t qo = (int64_t)g + c_half - 16;
                        uint32_t r = 0, ex = 0;
                        for (int t = 0; t < 32; ++t) {
                            const int64_t sp = ss + t;
                            if (sp < 0 || sp >= (int64_t)sl) continue;
                            const uint8_t wq = win[qo + t];
                            if (wq == QSENT) continue;
                            const uint8_t qr = red[qo + t];
                            if (qr == (sseed[sb + sp] & 0x7F)) {
                                ++r;
                                if (P->kx > 0 && wq == sres[sb + sp]) ++ex;
                            }
                        }
                        if (r + P->kx * ex < P->min_ids) {
                            ctr->hits_filtered++;
                            continue;
                        }
                    }
                    if (nsv == svcap) {
                        svcap = svcap ? 2 * svcap : 1024;
                        sv = (Surv*)xrealloc(sv, sizeof(Surv) * svcap);
                    }
                    sv[nsv].g = g;
                    sv[nsv].t = (uint32_t)p;
                    sv[nsv].shape = si;
                    ++nsv;
                }
            }
        }
        if (nsv == 0) continue;
        /* stable sort by batch position (insertion sort keeps the arrival
         * order (shape, subject position) among equal positions) */
        {
            Surv* tmp = (Surv*)malloc(sizeof(Surv) * nsv);
            /* merge sort, stable */
            for (uint64_t width = 1; width < nsv; width *= 2) {
                for (uint64_t i = 0; i < nsv; i += 2 * width) {
                    uint64_t a = i, am = i + width < nsv ? i + width : nsv;
                    uint64_t b = am, bm = i + 2 * width < nsv ? i + 2 * width : nsv, o = i;
                    while (a < am && b < bm) tmp[o++] = sv[b].g < sv[a].g ? sv[b++] : sv[a++];
                    while (a < am) tmp[o++] = sv[a++];
                    while (b < bm) tmp[o++] = sv[b++];
                }
                memcpy(sv, tmp, sizeof(Surv) * nsv);
            }
            free(tmp);
        }
        /* groups per query */
        uint64_t i = 0;
        while (i < nsv) {
            const uint32_t q = qof[sv[i].g];
            uint64_t j = i + 1;
            while (j < nsv && qof[sv[j].g] == q) ++j;
            ctr->scan_groups++;
            const uint64_t m = qoff[q + 1] - qoff[q];
            const uint8_t* qs = qres + qoff[q];
            const int8_t* qb = qbias ? qbias + qoff[q] : NULL;
            /* visit_scan: per-diagonal coverage, fresh for the group */
            const uint64_t nd = m + sl + 2;
            if (nd > dcap) {
                dcap = nd;
                dcov = (uint32_t*)xrealloc(dcov, sizeof(uint32_t) * dcap);
            }
            memset(dcov, 0, sizeof(uint32_t) * nd);
            if (j - i > hcap) {
                hcap = j - i;
                hsps = (Hsp*)xrealloc(hsps, sizeof(Hsp) * hcap);
            }
            uint64_t nh = 0;
            for (uint64_t k = i; k < j; ++k) {
                const uint32_t qp = sv[k].g - (uint32_t)base[q], sp = sv[k].t;
                if (log_survivors) surv_log(&res->log, q, s, qp, sp, sv[k].shape);
                const int64_t d = (int64_t)sp - (int64_t)qp;
                const uint64_t di = (uint64_t)(d + (int64_t)m);
                if (qp < dcov[di]) {
                    ctr->extensions_avoided++;
                    continue;
                }
                ctr->ungapped_extensions++;
                int32_t score;
                int64_t lo, hi;
                uint64_t steps = 0;
                or_xdrop(qs, (int64_t)m, qb, sres + sb, (int64_t)sl, mat, qp, sp, P->x, &score, &lo, &hi, &steps);
                ctr->xdrop_steps += steps;
                dcov[di] = (uint32_t)hi;
                if (score >= P->hsp_threshold) {
                    hsps[nh].diag = d;
                    hsps[nh].q_lo = (uint32_t)lo;
                    hsps[nh].q_hi = (uint32_t)hi;
                    hsps[nh].score = score;
                    ++nh;
                }
            }
            /* finish_subject_into */
            if (nh > 0) {
                ctr->hsps += nh;
                qsort(hsps, nh, sizeof(Hsp), cmp_hsp_diag);
                if (nh > rcap) {
                    rcap = nh;
                    regs = (or_region*)xrealloc(regs, sizeof(or_region) * rcap);
                }
                uint64_t nr = 0;
                const int64_t join = 2 * P->window_slack;
                for (uint64_t k = 0; k < nh; ++k) {
                    if (nr > 0 && hsps[k].diag - regs[nr - 1].d_max <= join) {
                        or_region* r = &regs[nr - 1];
                        if (hsps[k].diag > r->d_max) r->d_max = hsps[k].diag;
                        if (hsps[k].score > r->best) r->best = hsps[k].score;
                        r->n_hsp += 1;
                        const int64_t sum = (int64_t)r->hsp_sum + hsps[k].score;
                        r->hsp_sum = sum > INT32_MAX ? INT32_MAX : (sum < INT32_MIN ? INT32_MIN : (int32_t)sum);
                    } else {
                        or_region* r = &regs[nr++];
                        memset(r, 0, sizeof(*r));
                        r->subject = s;
                        r->d_min = r->d_max = hsps[k].diag;
                        r->best = hsps[k].score;
                        r->n_hsp = 1;
                        r->hsp_sum = hsps[k].score;
                    }
                }
                const int32_t thr = region_thr[q];
                if (thr > P->hsp_threshold) {
                    uint64_t kept = 0;
                    for (uint64_t k = 0; k < nr; ++k)
                        if (regs[k].best >= thr || (regs[k].n_hsp >= 2 && regs[k].hsp_sum >= thr)) regs[kept++] = regs[k];
                    ctr->regions_weak_dropped += nr - kept;
                    nr = kept;
                }
                qsort(regs, nr, sizeof(or_region), cmp_best_dmin);
                const uint64_t keep = nr < P->max_regions_per_subject ? nr : P->max_regions_per_subject;
                for (uint64_t k = 0; k < keep; ++k) region_push(&res->per_query[q], &regs[k]);
            }
            i = j;
        }
    }
    /* per-query cap under region_order */
    for (uint32_t q = 0; q < nq; ++q) {
        RegionList* R = &res->per_query[q];
        qsort(R->v, R->n, sizeof(or_region), cmp_region_order);
        if (R->n > P->cap) {
            ctr->regions_dropped_by_cap += R->n - P->cap;
            R->n = P->cap;
        }
        ctr->regions += R->n;
    }
    for (uint32_t si = 0; si < P->nshapes; ++si) free(tab[si]);
    free(tab);
    free(ntab);
    free(qof);
    free(sv);
    free(hsps);
    free(regs);
    free(dcov);
    free(sseed);
    free(red);
    free(win);
    free(base);
    (void)nres_q;
    return res;
}

uint64_t or_scan_nregions(void* h, uint32_t q) {
    const ScanResult* r = (const ScanResult*)h;
    return q < r->nq ? r->per_query[q].n : 0;
}

void or_scan_regions(void* h, uint32_t q, or_region* out) {
    const ScanResult* r = (const ScanResult*)h;
    if (q < r->nq && r->per_query[q].n) memcpy(out, r->per_query[q].v, sizeof(or_region) * r->per_query[q].n);
}

uint64_t or_scan_nsurvivors(void* h) { return ((const ScanResult*)h)->log.n; }

void or_scan_survivors(void* h, uint32_t* out) {
    const ScanResult* r = (const ScanResult*)h;
    if (r->log.n) memcpy(out, r->log.q, sizeof(uint32_t) * 5 * r->log.n);
}

void or_scan_free(void* h) {
    ScanResult* r = (ScanResult*)h;
    if (!r) return;
    for (uint32_t q = 0; q < r->nq; ++q) free(r->per_query[q].v);
    free(r->per_query);
    free(r->log.q);
    free(r);
}
