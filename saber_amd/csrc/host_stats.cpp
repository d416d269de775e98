// Karlin-Altschul statistics of the host pipeline: the thresholds the
// scoring path uses (two-stage band cut, region threshold, reporting cut and
// the learned prefilter's labels). Restates src/scoring/stats.rs:
// GapCosts::ncbi (:94-97), the NCBI parameter rows (:143-157, blast_stat.c),
// gapped_protein (:351-396), compute_length_adjustment (:458-521),
// SearchStatistics::for_query (:599-636) and the QueryStatistics methods
// (:638-760).
#include <cmath>
#include <cstdint>
#include <limits>

#include "../../include/saber_hip.h"

namespace {

struct Row {
    int existence, extension;
    double lambda, k, h, alpha, beta;
};

// NCBI blast_stat.c parameter rows (existence, extension, lambda, K, H,
// alpha, beta); the ungapped row first.
const Row kBlosum62[] = {
    {0, 0, 0.3176, 0.134, 0.4012, 0.7916, -3.2},
    {11, 2, 0.297, 0.082, 0.27, 1.1, -10.0}, {10, 2, 0.291, 0.075, 0.23, 1.3, -15.0},
    {9, 2, 0.279, 0.058, 0.19, 1.5, -19.0},  {8, 2, 0.264, 0.045, 0.15, 1.8, -26.0},
    {7, 2, 0.239, 0.027, 0.10, 2.5, -46.0},  {6, 2, 0.201, 0.012, 0.061, 3.3, -58.0},
    {13, 1, 0.292, 0.071, 0.23, 1.2, -11.0}, {12, 1, 0.283, 0.059, 0.19, 1.5, -19.0},
    {11, 1, 0.267, 0.041, 0.14, 1.9, -30.0}, {10, 1, 0.243, 0.024, 0.10, 2.5, -44.0},
    {9, 1, 0.206, 0.010, 0.052, 4.0, -87.0},
};
const Row kBlosum45[] = {
    {0, 0, 0.2291, 0.0924, 0.2514, 0.9113, -5.7},
    {13, 3, 0.207, 0.049, 0.14, 1.5, -22.0}, {12, 3, 0.199, 0.039, 0.11, 1.8, -34.0},
    {11, 3, 0.190, 0.031, 0.095, 2.0, -38.0}, {10, 3, 0.179, 0.023, 0.075, 2.4, -51.0},
    {18, 1, 0.207, 0.050, 0.15, 1.4, -28.0}, {17, 1, 0.198, 0.037, 0.12, 1.6, -33.0},
    {16, 1, 0.186, 0.025, 0.10, 1.9, -42.0}, {15, 1, 0.171, 0.015, 0.063, 2.7, -76.0},
};

void from_row(const Row& r, int calibrated, saber_hip_karlin* out) {
    out->lambda = r.lambda;
    out->k = r.k;
    out->log_k = std::log(r.k);
    out->h = r.h;
    out->alpha_d_lambda = r.alpha / r.lambda;
    out->beta = r.beta;
    out->calibrated = calibrated;
    out->even_scores_only = 0;
}

double adj_score(const saber_hip_query_stats* q, int64_t raw) {
    const int64_t s = (q->even_scores_only && raw % 2 != 0) ? raw - 1 : raw;
    return (double)s;
}

bool passes(const saber_hip_query_stats* q, int64_t s, double evalue_cut, double min_bits) {
    if (!(saber_hip_stats_evalue(q, s) <= evalue_cut)) return false;
    return std::isnan(min_bits) || saber_hip_stats_bitscore(q, s) >= min_bits;
}

}  // namespace

extern "C" {

int saber_hip_stats_gapped_protein(int matrix, int32_t gap_open, int32_t gap_extend, int allow_uncalibrated,
                                   saber_hip_karlin* out) {
    if (!out) return SABER_HIP_EINVAL;
    const Row* rows;
    int nrows;
    if (matrix == 62) {
        rows = kBlosum62;
        nrows = (int)(sizeof(kBlosum62) / sizeof(Row));
    } else if (matrix == 45) {
        rows = kBlosum45;
        nrows = (int)(sizeof(kBlosum45) / sizeof(Row));
    } else {
        return SABER_HIP_EINVAL;
    }
    // SABER costs (open = first residue) -> NCBI (existence, extension)
    const int a = gap_open - gap_extend, b = gap_extend;
    for (int i = 1; i < nrows; ++i)
        if (rows[i].existence == a && rows[i].extension == b) {
            from_row(rows[i], 1, out);
            return SABER_HIP_OK;
        }
    if (!allow_uncalibrated) return SABER_HIP_EINVAL;
    from_row(rows[0], 0, out);
    out->alpha_d_lambda = 1.0 / out->h;
    out->beta = 0.0;
    return SABER_HIP_OK;
}

int saber_hip_stats_length_adjustment(double k, double log_k, double alpha_d_lambda, double beta,
                                      int64_t query_length, int64_t db_length, int64_t db_num_seqs,
                                      int64_t* adj, int* converged) {
    const int kMaxIter = 20;
    const double m = (double)query_length, n = (double)db_length, nn = (double)db_num_seqs;
    double ell_min = 0.0, ell_max, ell_next = 0.0;
    bool conv = false;
    {
        const double a = nn, mb = m * nn + n, c = n * m - std::max(m, n) / k;
        if (c < 0.0) {
            if (adj) *adj = 0;
            if (converged) *converged = 0;
            return SABER_HIP_OK;
        }
        ell_max = 2.0 * c / (mb + std::sqrt(mb * mb - 4.0 * a * c));
    }
    for (int i = 1; i <= kMaxIter; ++i) {
        const double ell = ell_next;
        const double ss = (m - ell) * (n - nn * ell);
        const double ell_bar = alpha_d_lambda * (log_k + std::log(ss)) + beta;
        if (ell_bar >= ell) {
            ell_min = ell;
            if (ell_bar - ell_min <= 1.0) {
                conv = true;
                break;
            }
            if (ell_min == ell_max) break;
        } else {
            ell_max = ell;
        }
        if (ell_min <= ell_bar && ell_bar <= ell_max)
            ell_next = ell_bar;
        else
            ell_next = i == 1 ? ell_max : (ell_min + ell_max) / 2.0;
    }
    int64_t a = (int64_t)ell_min;
    if (conv) {
        const double ell = std::ceil(ell_min);
        if (ell <= ell_max) {
            const double ss = (m - ell) * (n - nn * ell);
            if (alpha_d_lambda * (log_k + std::log(ss)) + beta >= ell) a = (int64_t)ell;
        }
    }
    if (adj) *adj = a;
    if (converged) *converged = conv ? 1 : 0;
    return SABER_HIP_OK;
}

int saber_hip_stats_for_query(const saber_hip_karlin* p, uint64_t db_length, uint64_t db_num_seqs,
                              int per_subject, uint64_t query_len, saber_hip_query_stats* out) {
    if (!p || !out) return SABER_HIP_EINVAL;
    const uint64_t dbl = db_length ? db_length : 1, dbn = db_num_seqs ? db_num_seqs : 1;
    const int64_t m = query_len ? (int64_t)query_len : 1;
    int64_t adj = 0;
    saber_hip_stats_length_adjustment(p->k, p->log_k, p->alpha_d_lambda, p->beta, m, (int64_t)dbl, (int64_t)dbn,
                                      &adj, nullptr);
    const int64_t min_q = (int64_t)std::ceil(1.0 / p->k);
    if (m - adj < min_q) adj = std::max<int64_t>(m - min_q, 0);
    const double eff_q = std::max((double)(m - adj), 1.0);
    const double eff_db = std::max((double)dbl - (double)dbn * (double)adj, 1.0);
    out->lambda = p->lambda;
    out->k = p->k;
    out->log_k = p->log_k;
    out->length_adjustment = adj;
    out->eff_query_len = eff_q;
    out->eff_db_len = eff_db;
    out->search_space = eff_q * eff_db;
    out->even_scores_only = p->even_scores_only;
    out->per_subject = per_subject ? 1 : 0;
    out->query_len = m;
    out->num_seqs = (double)dbn;
    out->alpha_d_lambda = p->alpha_d_lambda;
    out->beta = p->beta;
    return SABER_HIP_OK;
}

double saber_hip_stats_space_for(const saber_hip_query_stats* q, uint64_t subject_len) {
    if (!q->per_subject) return q->search_space;
    const int64_t m = q->query_len > 0 ? q->query_len : 1;
    const int64_t n = subject_len ? (int64_t)subject_len : 1;
    int64_t adj = 0;
    saber_hip_stats_length_adjustment(q->k, q->log_k, q->alpha_d_lambda, q->beta, m, n, 1, &adj, nullptr);
    const int64_t min_l = (int64_t)std::ceil(1.0 / q->k);
    adj = std::max<int64_t>(std::min(std::min(adj, std::max<int64_t>(m - min_l, 0)), std::max<int64_t>(n - min_l, 0)), 0);
    return q->num_seqs * (double)std::max<int64_t>(m - adj, 1) * (double)std::max<int64_t>(n - adj, 1);
}

double saber_hip_stats_space_lower_bound(const saber_hip_query_stats* q, uint64_t subject_len) {
    if (!q->per_subject) return q->search_space;
    const double m = (double)(q->query_len > 0 ? q->query_len : 1);
    const double n = (double)(subject_len ? subject_len : 1);
    const double up = std::ceil(std::max(q->alpha_d_lambda * (q->log_k + std::log(m * n)) + q->beta, 0.0));
    return q->num_seqs * std::max(m - up, 1.0) * std::max(n - up, 1.0);
}

double saber_hip_stats_floor_lower_bound(const saber_hip_query_stats* q, double evalue_cut, double min_bits,
                                         uint64_t subject_len, int64_t db_floor) {
    if (!q->per_subject) return (double)db_floor;
    double f = 1.0;
    if (evalue_cut > 0.0 && std::isfinite(evalue_cut))
        f = std::max(f, std::log(q->k * saber_hip_stats_space_lower_bound(q, subject_len) / evalue_cut) / q->lambda);
    if (!std::isnan(min_bits)) f = std::max(f, (min_bits * M_LN2 + q->log_k) / q->lambda);
    if (q->even_scores_only) f -= 1.0;
    return f;
}

double saber_hip_stats_evalue_in(const saber_hip_query_stats* q, int64_t raw, double space) {
    if (raw <= 0) return std::numeric_limits<double>::infinity();
    const double e = q->k * space * std::exp(-q->lambda * adj_score(q, raw));
    return e == 0.0 ? std::numeric_limits<double>::min() : e;
}

double saber_hip_stats_evalue_for(const saber_hip_query_stats* q, int64_t raw, uint64_t subject_len) {
    return saber_hip_stats_evalue_in(q, raw, saber_hip_stats_space_for(q, subject_len));
}

double saber_hip_stats_evalue(const saber_hip_query_stats* q, int64_t raw) {
    return saber_hip_stats_evalue_in(q, raw, q->search_space);
}

double saber_hip_stats_bitscore(const saber_hip_query_stats* q, int64_t raw) {
    if (raw <= 0) return 0.0;
    return (q->lambda * adj_score(q, raw) - q->log_k) / M_LN2;
}

int64_t saber_hip_stats_min_score(const saber_hip_query_stats* q, double evalue_cut, double min_bits) {
    double floor = 1.0;
    if (evalue_cut > 0.0 && std::isfinite(evalue_cut))
        floor = std::max(floor, std::log((q->k * q->search_space) / evalue_cut) / q->lambda);
    if (!std::isnan(min_bits)) floor = std::max(floor, (min_bits * M_LN2 + q->log_k) / q->lambda);
    if (!std::isfinite(floor)) return 1;
    int64_t s = (int64_t)std::max(std::ceil(floor), 1.0);
    while (s > 1 && passes(q, s - 1, evalue_cut, min_bits)) --s;
    while (!passes(q, s, evalue_cut, min_bits) && s < std::numeric_limits<int64_t>::max() / 2) ++s;
    return s;
}

}  // extern "C"

extern "C" int saber_hip_stats_evalue_batch(const saber_hip_karlin* p, uint64_t db_length, uint64_t db_num_seqs,
                                            int per_subject, uint64_t n, const uint64_t* query_len,
                                            const uint64_t* subject_len, const int64_t* raw, double* evalue) {
    if (!p || (n && (!query_len || !subject_len || !raw || !evalue))) return SABER_HIP_EINVAL;
    saber_hip_query_stats q{};
    uint64_t last_m = ~0ull;
    for (uint64_t i = 0; i < n; ++i) {
        if (query_len[i] != last_m) {
            saber_hip_stats_for_query(p, db_length, db_num_seqs, per_subject, query_len[i], &q);
            last_m = query_len[i];
        }
        evalue[i] = saber_hip_stats_evalue_for(&q, raw[i], subject_len[i]);
    }
    return SABER_HIP_OK;
}
