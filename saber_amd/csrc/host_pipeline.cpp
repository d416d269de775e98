// Host pipeline around the scoring kernels: hit visiting with diagonal
// coverage (visit_scan, src/search/engine/scan.rs:2384-2431), HSP -> region
// joining (finish_subject, src/search/engine.rs:2350-2412 and
// finish_subject_into, scan.rs:2437-2474), the per-query region cap
// (region_order, engine.rs:493-495) and window / band planning (window and
// plan_windows, engine.rs:1223-1324; Band, src/algorithm/banded.rs:30-54).
#include <algorithm>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/saber_hip.h"
#include "plan_rules.h"

namespace {

using saber_plan::Band;
using saber_plan::kBandBridge;
using saber_plan::window;

int32_t sat_add(int32_t a, int32_t b) {
    const int64_t s = (int64_t)a + b;
    return (int32_t)std::max<int64_t>(std::min<int64_t>(s, std::numeric_limits<int32_t>::max()),
                                      std::numeric_limits<int32_t>::min());
}

}  // namespace

extern "C" {

int saber_hip_visit_hits(uint32_t qlen, uint32_t slen, const uint32_t* qp, const uint32_t* sp, uint64_t n,
                         const int32_t* score, const uint32_t* q_lo, const uint32_t* q_hi, int32_t hsp_threshold,
                         saber_hip_hsp* hsps, uint64_t* nhsp, uint64_t* extended, uint64_t* avoided) {
    if (n && (!qp || !sp || !score || !q_lo || !q_hi || !hsps)) return SABER_HIP_EINVAL;
    // Diagonal d = sp - qp in [-qlen, slen]: cover[d + qlen] = q_hi of the last
    // extension on it (0 = none yet).
    std::vector<uint32_t> cover((size_t)qlen + slen + 2, 0);
    uint64_t nh = 0, ext = 0, av = 0;
    for (uint64_t k = 0; k < n; ++k) {
        const int64_t d = (int64_t)sp[k] - (int64_t)qp[k];
        const int64_t di = d + (int64_t)qlen;
        if (di < 0 || di >= (int64_t)cover.size() || sp[k] >= slen || qp[k] >= qlen) continue;
        if (qp[k] < cover[(size_t)di]) {
            ++av;
            continue;
        }
        ++ext;
        cover[(size_t)di] = q_hi[k];
        if (score[k] >= hsp_threshold) hsps[nh++] = saber_hip_hsp{d, q_lo[k], q_hi[k], score[k], 0};
    }
    if (nhsp) *nhsp = nh;
    if (extended) *extended = ext;
    if (avoided) *avoided = av;
    return SABER_HIP_OK;
}

int saber_hip_finish_subject(uint64_t subject, saber_hip_hsp* hsps, uint64_t n, int64_t window_slack,
                             int32_t hsp_threshold, int32_t region_thr, uint32_t max_regions_per_subject,
                             saber_hip_region* out, uint32_t* nout) {
    if (!nout || (n && (!hsps || !out))) return SABER_HIP_EINVAL;
    *nout = 0;
    if (n == 0) return SABER_HIP_OK;
    std::sort(hsps, hsps + n, [](const saber_hip_hsp& a, const saber_hip_hsp& b) { return a.diag < b.diag; });
    const int64_t join = 2 * window_slack;
    std::vector<saber_hip_region> regs;
    for (uint64_t k = 0; k < n; ++k) {
        const saber_hip_hsp& h = hsps[k];
        if (!regs.empty() && h.diag - regs.back().d_max <= join) {
            saber_hip_region& r = regs.back();
            r.d_max = std::max(r.d_max, h.diag);
            r.best = std::max(r.best, h.score);
            r.n_hsp += 1;
            r.hsp_sum = sat_add(r.hsp_sum, h.score);
        } else {
            regs.push_back(saber_hip_region{subject, h.diag, h.diag, h.score, 1, h.score});
        }
    }
    if (region_thr > hsp_threshold) {
        regs.erase(std::remove_if(regs.begin(), regs.end(),
                                  [&](const saber_hip_region& r) {
                                      return !(r.best >= region_thr || (r.n_hsp >= 2 && r.hsp_sum >= region_thr));
                                  }),
                   regs.end());
    }
    std::sort(regs.begin(), regs.end(), [](const saber_hip_region& a, const saber_hip_region& b) {
        return a.best != b.best ? a.best > b.best : a.d_min < b.d_min;
    });
    const size_t keep = std::min<size_t>(regs.size(), std::max<uint32_t>(max_regions_per_subject, 1));
    std::copy(regs.begin(), regs.begin() + keep, out);
    *nout = (uint32_t)keep;
    return SABER_HIP_OK;
}

int saber_hip_cap_regions(saber_hip_region* regs, uint64_t n, uint64_t cap, uint64_t* nout) {
    if (!nout || (n && !regs)) return SABER_HIP_EINVAL;
    std::sort(regs, regs + n, [](const saber_hip_region& a, const saber_hip_region& b) {
        if (a.best != b.best) return a.best > b.best;
        if (a.subject != b.subject) return a.subject < b.subject;
        if (a.d_min != b.d_min) return a.d_min < b.d_min;
        return a.d_max < b.d_max;
    });
    *nout = std::min(n, cap);
    return SABER_HIP_OK;
}

int saber_hip_plan_windows(const saber_hip_region* regs, uint64_t n, uint64_t qlen, const uint64_t* subject_len,
                           const saber_hip_plan_params* p, saber_hip_plan* out) {
    if (!p || (n && (!regs || !subject_len || !out))) return SABER_HIP_EINVAL;
    const int64_t q = (int64_t)qlen;
    // Chains of strong regions per subject (single linkage within kBandBridge).
    // The regions are grouped by subject through one index sort (subject,
    // then region index), without a per-subject container.
    std::vector<std::pair<int64_t, int64_t>> span(n);
    std::vector<uint64_t> order(n), idx;
    for (uint64_t i = 0; i < n; ++i) {
        span[i] = {regs[i].d_min, regs[i].d_max};
        order[i] = i;
    }
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
        return regs[a].subject != regs[b].subject ? regs[a].subject < regs[b].subject : a < b;
    });
    for (uint64_t g0 = 0, g1 = 0; g0 < n; g0 = g1) {
        g1 = g0 + 1;
        while (g1 < n && regs[order[g1]].subject == regs[order[g0]].subject) ++g1;
        int32_t top = 0;
        for (uint64_t k = g0; k < g1; ++k) top = std::max(top, regs[order[k]].best);
        const int32_t strong = saber_plan::strong_cut(p->hsp_threshold, top);
        idx.clear();
        for (uint64_t k = g0; k < g1; ++k)
            if (regs[order[k]].best >= strong) idx.push_back(order[k]);
        std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
            if (regs[a].d_min != regs[b].d_min) return regs[a].d_min < regs[b].d_min;
            if (regs[a].d_max != regs[b].d_max) return regs[a].d_max < regs[b].d_max;
            return a < b;
        });
        size_t start = 0;
        while (start < idx.size()) {
            int64_t lo = regs[idx[start]].d_min, hi = regs[idx[start]].d_max;
            size_t end = start + 1;
            while (end < idx.size() && regs[idx[end]].d_min - hi <= kBandBridge) {
                hi = std::max(hi, regs[idx[end]].d_max);
                ++end;
            }
            for (size_t k = start; k < end; ++k) span[idx[k]] = {lo, hi};
            start = end;
        }
    }
    for (uint64_t ri = 0; ri < n; ++ri) {
        const saber_hip_region& r = regs[ri];
        out[ri] = saber_plan::plan_region(r.d_min, r.d_max, r.best, span[ri].first, span[ri].second, q,
                                          (int64_t)subject_len[ri], *p);
    }
    return SABER_HIP_OK;
}


// composition_bias (src/scoring/bias.rs:56-122): the per-position correction
// c_i = sum_a p_a S(q_i, a) - (sum over the +-20 window, j != i, of
// S(q_i, q_j)) / (|W| - 1), rounded half away from zero, clamped to +-127;
// 0 for non-standard residues and for queries shorter than 2. Floating-point
// operations in the reference's order (Robinson frequencies normalised by
// their left-to-right sum, background scores summed a = 0..19), so every
// rounding is the reference's.
int saber_hip_composition_bias(const uint8_t* query, uint64_t n, const int32_t* matrix, int8_t* out) {
    if (n > 0 && (!query || !out)) return 1;
    if (!matrix) return 1;
    for (uint64_t i = 0; i < n; ++i) out[i] = 0;
    if (n < 2) return 0;
    static const double kRobinson[20] = {78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95,
                                         73.77, 21.99, 51.42, 90.19, 57.44, 22.43, 38.56,
                                         52.03, 71.20, 58.41, 13.30, 32.16, 64.41};
    double tot = 0.0;
    for (double v : kRobinson) tot += v;
    double p[20];
    for (int a = 0; a < 20; ++a) p[a] = kRobinson[a] / tot;
    double bg[20];
    for (int r = 0; r < 20; ++r) {
        double acc = 0.0;
        for (int a = 0; a < 20; ++a) acc += p[a] * (double)matrix[r * 24 + a];
        bg[r] = acc;
    }
    int32_t acc[20] = {0};
    const uint64_t half = 20;
    uint64_t lo = 0, hi = 0;   // window [lo, hi)
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t want_hi = std::min(i + half + 1, n);
        const uint64_t want_lo = i > half ? i - half : 0;
        while (hi < want_hi) {
            const int c = std::min<int>(query[hi], 23);
            for (int r = 0; r < 20; ++r) acc[r] += matrix[r * 24 + c];
            ++hi;
        }
        while (lo < want_lo) {
            const int c = std::min<int>(query[lo], 23);
            for (int r = 0; r < 20; ++r) acc[r] -= matrix[r * 24 + c];
            ++lo;
        }
        const int r = query[i];
        if (r < 20) {
            const int32_t w = (int32_t)(hi - lo);
            if (w >= 2) {
                const int32_t own = matrix[r * 24 + r];
                const double neigh = (double)(acc[r] - own) / (double)(w - 1);
                const double f = bg[r] - neigh;
                const int32_t v = (int32_t)(f < 0.0 ? f - 0.5 : f + 0.5);
                out[i] = (int8_t)std::max(-127, std::min(127, v));
            }
        }
    }
    return 0;
}

}  // extern "C"
