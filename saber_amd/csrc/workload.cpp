// Synthetic workloads (host only, deterministic from a seed): protein
// sequences drawn i.i.d. from the Robinson & Robinson background
// (src/scoring/bias.rs:32-53), lognormal lengths, and planted homologues made
// by BLOSUM62-conditional substitutions at a target identity plus geometric
// indels. Used by the benchmark and the tests to build the banded-region and
// full-SW batches; nothing here is on the scoring path.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double normal() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

const double kRobinson[20] = {78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95,
                              73.77, 21.99, 51.42, 90.19, 57.44, 22.43, 38.56,
                              52.03, 71.20, 58.41, 13.30, 32.16, 64.41};

struct Model {
    double cdf[20];
    double cond[20][20];  // cumulative P(b | a, b != a) from the matrix target frequencies
    Model(const int32_t* mat) {
        double tot = 0;
        for (double v : kRobinson) tot += v;
        double acc = 0;
        for (int a = 0; a < 20; ++a) {
            acc += kRobinson[a] / tot;
            cdf[a] = acc;
        }
        const double lambda = 0.3176;  // ungapped BLOSUM62 lambda with these frequencies
        for (int a = 0; a < 20; ++a) {
            double w[20], sum = 0;
            for (int b = 0; b < 20; ++b) {
                w[b] = b == a ? 0.0 : (kRobinson[b] / tot) * std::exp(lambda * mat[a * 24 + b]);
                sum += w[b];
            }
            double c = 0;
            for (int b = 0; b < 20; ++b) {
                c += w[b] / sum;
                cond[a][b] = c;
            }
        }
    }
    uint8_t residue(Rng& r) const {
        double u = r.uni();
        for (int a = 0; a < 20; ++a)
            if (u < cdf[a]) return (uint8_t)a;
        return 19;
    }
    uint8_t substitute(Rng& r, uint8_t a) const {
        double u = r.uni();
        for (int b = 0; b < 20; ++b)
            if (u < cond[a][b]) return (uint8_t)b;
        return (uint8_t)(a == 19 ? 18 : 19);
    }
};

uint32_t lognormal_len(Rng& r, double median, double sigma, uint32_t lo, uint32_t hi) {
    double v = median * std::exp(sigma * r.normal());
    long x = std::lround(v);
    if (x < (long)lo) x = lo;
    if (x > (long)hi) x = hi;
    return (uint32_t)x;
}

uint32_t geometric(Rng& r, double mean) {
    // P(len = k) = p (1-p)^(k-1), k >= 1, mean 1/p
    const double p = 1.0 / mean;
    uint32_t k = 1;
    while (r.uni() >= p && k < 64) ++k;
    return k;
}

// Mutate `src` into `out` at the given identity with geometric indels; track
// the diagonal drift of the planted alignment.
void mutate(const Model& M, Rng& r, const uint8_t* src, uint32_t len, double identity,
            double indel_rate, double indel_mean, std::vector<uint8_t>& out) {
    for (uint32_t i = 0; i < len;) {
        if (r.uni() < indel_rate) {
            const uint32_t k = geometric(r, indel_mean);
            if (r.below(2) == 0) {
                for (uint32_t t = 0; t < k; ++t) out.push_back(M.residue(r));
            } else {
                i += k;
                continue;
            }
        }
        out.push_back(r.uni() < identity ? src[i] : M.substitute(r, src[i]));
        ++i;
    }
}

struct Batch {
    std::vector<uint8_t> qres, sres;
    std::vector<uint64_t> qoff{0}, soff{0};
    std::vector<int32_t> band_lo, band_hi;
    std::vector<uint8_t> homolog;
    std::vector<int32_t> source;
};

}  // namespace

extern "C" {

// Banded-region batch: n regions over ceil(n / per_query) queries (the
// engine scores all band jobs of one query together). Region k pairs query
// k / per_query with its own subject window and band [c-h, c+h]: a planted
// homologue (mutated copy of the query with random flanks, c = left flank)
// or an unrelated window (c uniform over the middle of the rectangle).
void* wl_band_regions(uint64_t seed, uint64_t n, uint32_t per_query, double median, double sigma,
                      uint32_t lmin, uint32_t lmax, double homolog_frac,
                      const int32_t* half_widths, int n_half, const int32_t* mat) {
    auto* B = new Batch();
    Model M(mat);
    Rng r(seed);
    std::vector<uint8_t> tmp;
    if (per_query == 0) per_query = 1;
    size_t q0 = 0;
    uint32_t m = 0;
    for (uint64_t k = 0; k < n; ++k) {
        if (k % per_query == 0) {
            m = lognormal_len(r, median, sigma, lmin, lmax);
            q0 = B->qres.size();
            for (uint32_t i = 0; i < m; ++i) B->qres.push_back(M.residue(r));
            B->qoff.push_back(B->qres.size());
        }
        const int32_t h = half_widths[r.below((uint64_t)n_half)];
        const bool hom = r.uni() < homolog_frac;
        int64_t c;
        if (hom) {
            const double ident = 0.25 + 0.35 * r.uni();
            const uint32_t lf = (uint32_t)r.below(31), rf = (uint32_t)r.below(31);
            for (uint32_t t = 0; t < lf; ++t) B->sres.push_back(M.residue(r));
            tmp.clear();
            mutate(M, r, B->qres.data() + q0, m, ident, 0.02, 3.0, tmp);
            B->sres.insert(B->sres.end(), tmp.begin(), tmp.end());
            for (uint32_t t = 0; t < rf; ++t) B->sres.push_back(M.residue(r));
            c = (int64_t)lf;
        } else {
            const uint32_t nl = lognormal_len(r, median, sigma, lmin, lmax);
            for (uint32_t t = 0; t < nl; ++t) B->sres.push_back(M.residue(r));
            const int64_t span = (int64_t)m / 2 + nl / 2 + 1;
            c = (int64_t)r.below((uint64_t)span) - (int64_t)m / 2;
        }
        B->soff.push_back(B->sres.size());
        B->band_lo.push_back((int32_t)(c - h));
        B->band_hi.push_back((int32_t)(c + h));
        B->homolog.push_back(hom ? 1 : 0);
    }
    return B;
}

// Sequence set: n sequences with lognormal lengths (used for queries and
// subjects of the full-SW batch).
void* wl_sequences(uint64_t seed, uint64_t n, double median, double sigma, uint32_t lmin,
                   uint32_t lmax, const int32_t* mat) {
    auto* B = new Batch();
    Model M(mat);
    Rng r(seed);
    for (uint64_t k = 0; k < n; ++k) {
        const uint32_t m = lognormal_len(r, median, sigma, lmin, lmax);
        for (uint32_t i = 0; i < m; ++i) B->qres.push_back(M.residue(r));
        B->qoff.push_back(B->qres.size());
    }
    return B;
}

// Search set (blastp workloads): nq queries with lognormal lengths; nsub
// subjects, each with probability homolog_frac a planted homologue of a
// uniformly chosen query (the whole query mutated at an identity uniform in
// [id_lo, id_hi] with geometric indels 0.02/residue, mean 3, between random
// flanks of 0..30 residues), else an unrelated sequence of lognormal length.
// source[k] = query index + 1 of subject k's template (0 = unrelated).
void* wl_search_set(uint64_t seed, uint64_t nq, uint64_t nsub, double median, double sigma, uint32_t lmin,
                    uint32_t lmax, double homolog_frac, double id_lo, double id_hi, const int32_t* mat) {
    auto* B = new Batch();
    Model M(mat);
    Rng r(seed);
    for (uint64_t k = 0; k < nq; ++k) {
        const uint32_t m = lognormal_len(r, median, sigma, lmin, lmax);
        for (uint32_t i = 0; i < m; ++i) B->qres.push_back(M.residue(r));
        B->qoff.push_back(B->qres.size());
    }
    std::vector<uint8_t> tmp;
    for (uint64_t k = 0; k < nsub; ++k) {
        if (nq > 0 && r.uni() < homolog_frac) {
            const uint64_t q = r.below(nq);
            const double ident = id_lo + (id_hi - id_lo) * r.uni();
            const uint32_t lf = (uint32_t)r.below(31), rf = (uint32_t)r.below(31);
            for (uint32_t t = 0; t < lf; ++t) B->sres.push_back(M.residue(r));
            tmp.clear();
            mutate(M, r, B->qres.data() + B->qoff[q], (uint32_t)(B->qoff[q + 1] - B->qoff[q]), ident, 0.02, 3.0,
                   tmp);
            B->sres.insert(B->sres.end(), tmp.begin(), tmp.end());
            for (uint32_t t = 0; t < rf; ++t) B->sres.push_back(M.residue(r));
            B->source.push_back((int32_t)(q + 1));
        } else {
            const uint32_t n = lognormal_len(r, median, sigma, lmin, lmax);
            for (uint32_t t = 0; t < n; ++t) B->sres.push_back(M.residue(r));
            B->source.push_back(0);
        }
        B->soff.push_back(B->sres.size());
    }
    return B;
}

// Family set (the self-search at scale): nq queries in planted families
// (a family's root has lognormal length; its members are mutated copies of
// the root at an identity uniform in [id_lo, id_hi] with geometric indels;
// sizes 2 + geometric(mean fam_extra), at most fam_max, the last family cut
// to fill nq), in shuffled order; the database is the queries followed by
// ndecoy reversed unrelated sequences of lognormal length (decoy_median,
// decoy_sigma, clamped to [lmin, decoy_max]). source[k] = family id of
// query k.
void* wl_family_set(uint64_t seed, uint64_t nq, double median, double sigma, uint32_t lmin, uint32_t lmax,
                    double fam_extra, uint32_t fam_max, double id_lo, double id_hi, uint64_t ndecoy,
                    double decoy_median, double decoy_sigma, uint32_t decoy_max, const int32_t* mat) {
    auto* B = new Batch();
    Model M(mat);
    Rng r(seed);
    std::vector<std::vector<uint8_t>> members;
    std::vector<int32_t> fam;
    std::vector<uint8_t> root, tmp;
    int32_t f = 0;
    while (members.size() < nq) {
        uint32_t size = 2;
        const double p = 1.0 / (1.0 + fam_extra);
        while (size < fam_max && r.uni() >= p) ++size;
        const uint32_t m = lognormal_len(r, median, sigma, lmin, lmax);
        root.clear();
        for (uint32_t i = 0; i < m; ++i) root.push_back(M.residue(r));
        for (uint32_t k = 0; k < size && members.size() < nq; ++k) {
            tmp.clear();
            mutate(M, r, root.data(), m, id_lo + (id_hi - id_lo) * r.uni(), 0.02, 3.0, tmp);
            if (tmp.empty()) tmp.push_back(M.residue(r));
            members.push_back(tmp);
            fam.push_back(f);
        }
        ++f;
    }
    // shuffle (Fisher-Yates)
    std::vector<uint64_t> order(members.size());
    for (uint64_t i = 0; i < order.size(); ++i) order[i] = i;
    for (uint64_t i = order.size(); i > 1; --i) std::swap(order[i - 1], order[r.below(i)]);
    for (uint64_t i : order) {
        B->qres.insert(B->qres.end(), members[i].begin(), members[i].end());
        B->qoff.push_back(B->qres.size());
        B->source.push_back(fam[i]);
    }
    B->sres = B->qres;
    B->soff = B->qoff;
    for (uint64_t k = 0; k < ndecoy; ++k) {
        const uint32_t n = lognormal_len(r, decoy_median, decoy_sigma, lmin, decoy_max);
        tmp.clear();
        for (uint32_t t = 0; t < n; ++t) tmp.push_back(M.residue(r));
        B->sres.insert(B->sres.end(), tmp.rbegin(), tmp.rend());
        B->soff.push_back(B->sres.size());
    }
    return B;
}

void wl_source(void* h, int32_t* out) {
    auto* B = static_cast<Batch*>(h);
    if (out && !B->source.empty()) std::memcpy(out, B->source.data(), 4 * B->source.size());
}

void wl_sizes(void* h, uint64_t* nqres, uint64_t* nq, uint64_t* nsres, uint64_t* ns) {
    auto* B = static_cast<Batch*>(h);
    *nqres = B->qres.size();
    *nq = B->qoff.size() - 1;
    *nsres = B->sres.size();
    *ns = B->soff.size() - 1;
}

void wl_copy(void* h, uint8_t* qres, uint64_t* qoff, uint8_t* sres, uint64_t* soff,
             int32_t* band_lo, int32_t* band_hi, uint8_t* homolog) {
    auto* B = static_cast<Batch*>(h);
    if (qres && !B->qres.empty()) std::memcpy(qres, B->qres.data(), B->qres.size());
    if (qoff) std::memcpy(qoff, B->qoff.data(), 8 * B->qoff.size());
    if (sres && !B->sres.empty()) std::memcpy(sres, B->sres.data(), B->sres.size());
    if (soff) std::memcpy(soff, B->soff.data(), 8 * B->soff.size());
    if (band_lo && !B->band_lo.empty()) std::memcpy(band_lo, B->band_lo.data(), 4 * B->band_lo.size());
    if (band_hi && !B->band_hi.empty()) std::memcpy(band_hi, B->band_hi.data(), 4 * B->band_hi.size());
    if (homolog && !B->homolog.empty()) std::memcpy(homolog, B->homolog.data(), B->homolog.size());
}

void wl_free(void* h) { delete static_cast<Batch*>(h); }

}  // extern "C"
