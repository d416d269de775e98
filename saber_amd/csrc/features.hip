// Feature vectors of candidate regions for the learned prefilter, computed on
// the device from what the pipeline holds when the prefilter runs (after the
// X-drop stage and the window/band plan, before banded scoring): the region's
// X-drop evidence, its geometry and the residue composition of the query and
// the subject window. Query compositions are computed once per query; then
// one lane per region.
//
// Features (bf16, 64 per region):
//   0  best X-drop score of the region's hits / 64
//   1  ln(1 + max(best, 0))
//   2  hits with score >= hsp_threshold (HSPs) / hits per region
//   3  sum of the HSP scores / 128
//   4  diagonal spread of the HSPs (max - min sp - qp) / 64
//   5  clipped band width / 128
//   6  ln(1 + window length)         7  ln(1 + query length)
//   8  best / ln(2 + query length x window length)
//   9  mean X-drop score of the hits / 64
//   10 longest extension (q_hi - q_lo) / 64
//   11 best / (1 + longest extension)
//   12..31 query composition (fractions of codes 0..19)
//   32..51 subject-window composition
//   52..63 zero
// Optionally standardised on the device, (f - mean[i]) * inv_std[i], the
// form the prefilter model is trained on.
#include "../../include/saber_hip.h"
#include "common.h"

namespace saber {

// Query composition, once per query (regions of one query share it): counts
// of codes 0..19 over the query / its length. One wave per query.
__global__ __launch_bounds__(256) void query_composition_kernel(const uint64_t* qoff, const uint8_t* qres,
                                                                uint32_t nq, float* comp) {
    __shared__ uint32_t hist[4][24];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t qi = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; qi < nq; qi += nw) {
        if (lane < 24) hist[wv][lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint8_t* q = qres + qoff[qi];
        const uint64_t m = qoff[qi + 1] - qoff[qi];
        for (uint64_t i = lane; i < m; i += 64) atomicAdd(&hist[wv][min((uint32_t)q[i], 23u)], 1u);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < 20) comp[(uint64_t)qi * 20 + lane] = m ? (float)hist[wv][lane] / (float)m : 0.f;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// Subject-window composition into H[0..23]: aligned 16-byte blocks from the
// one holding the window's first residue; loads past the window's last
// block re-read that block, and bytes outside the window are not counted
// (every block read lies inside the buffer, guard bytes included).
__device__ __forceinline__ void window_hist(const uint8_t* s, uint64_t nwin, uint32_t* H) {
#pragma unroll
    for (int c = 0; c < 24; ++c) H[c] = 0;
    if (!nwin) return;
    const uintptr_t p = reinterpret_cast<uintptr_t>(s), pe = p + nwin;
    const uintptr_t first = p & ~uintptr_t(15), last = (pe - 1) & ~uintptr_t(15);
    const uint8_t* s0 = s - (p & 15);   // the first block, as a pointer
    for (uintptr_t base = first; base <= last; base += 128) {
        uint4 v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uintptr_t at = base + 16 * j;
            const uintptr_t la = at <= last ? at : last;
            v[j] = *reinterpret_cast<const uint4*>(s0 + (la - first));
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uintptr_t at = base + 16 * j;
            const int64_t b_lo = (int64_t)p - (int64_t)at;   // count bytes b with b_lo <= b < b_hi
            const int64_t b_hi = (int64_t)pe - (int64_t)at;
            const uint32_t d[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (b >= b_lo && b < b_hi) atomicAdd(&H[min((d[b >> 2] >> (8 * (b & 3))) & 0xFFu, 23u)], 1u);
        }
    }
}

// Standardise (optional) and store 64 features as bf16, eight 16-byte rows.
__device__ __forceinline__ void store_features(float* f, const float* mean, const float* inv_std, __bf16* dst) {
    if (mean) {
#pragma unroll
        for (int c = 0; c < 64; ++c) f[c] = (f[c] - mean[c]) * inv_std[c];
    }
    uint4* out = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        uint32_t w4[4];
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const __bf16 lo = (__bf16)f[8 * x + 2 * y], hi = (__bf16)f[8 * x + 2 * y + 1];
            uint16_t bl, bh;
            __builtin_memcpy(&bl, &lo, 2);
            __builtin_memcpy(&bh, &hi, 2);
            w4[y] = (uint32_t)bl | ((uint32_t)bh << 16);
        }
        out[x] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
}

// One lane per region: the region's X-drop evidence, the composition of its
// subject window (a per-lane histogram in LDS), its geometry and its query's
// composition; the 64 features are stored as eight 16-byte rows. The kernel
// is bound by memory latency, so the window is read in rounds of eight
// independent 16-byte loads (128 residues per round trip).
constexpr int kHistStride = 25;   // 24 codes, odd stride against LDS bank conflicts

__global__ __launch_bounds__(256) void region_features_kernel(FeatureArgs a) {
    __shared__ uint32_t hist[256 * kHistStride];
    uint32_t* H = hist + threadIdx.x * kHistStride;
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    const saber_hip_band_job J = a.jobs[r];
    const uint64_t m = a.qoff[J.query + 1] - a.qoff[J.query];
    const uint8_t* s = a.sres + a.soff[J.subject] + J.win_lo;
    const uint64_t nwin = J.win_hi - J.win_lo;

    window_hist(s, nwin, H);

    // X-drop evidence of the region's hits (four hits per 16-byte load when
    // the region's hits start 16-byte aligned)
    int32_t best = a.per_region ? INT32_MIN : 0, hsp = 0, hsum = 0, ext = 0;
    int64_t dmin = INT64_MAX, dmax = INT64_MIN;
    float ssum = 0.f;
    auto take = [&](int32_t sc, uint32_t lo, uint32_t hi, uint32_t qp, uint32_t sp) {
        best = max(best, sc);
        ssum += (float)sc;
        ext = max(ext, (int32_t)(hi - lo));
        if (sc >= a.hsp_threshold) {
            ++hsp;
            hsum += sc;
            const int64_t d = (int64_t)sp - (int64_t)qp;
            dmin = min(dmin, d);
            dmax = max(dmax, d);
        }
    };
    const uint64_t h0 = r * a.per_region;
    if ((a.per_region & 3) == 0) {
        for (uint32_t k = 0; k < a.per_region; k += 4) {
            const uint64_t h = h0 + k;
            const int4 sc = *reinterpret_cast<const int4*>(a.xscore + h);
            const uint4 lo = *reinterpret_cast<const uint4*>(a.xlo + h);
            const uint4 hi = *reinterpret_cast<const uint4*>(a.xhi + h);
            const uint4 t0 = *reinterpret_cast<const uint4*>(a.hits + 4 * h);
            const uint4 t1 = *reinterpret_cast<const uint4*>(a.hits + 4 * h + 4);
            const uint4 t2 = *reinterpret_cast<const uint4*>(a.hits + 4 * h + 8);
            const uint4 t3 = *reinterpret_cast<const uint4*>(a.hits + 4 * h + 12);
            take(sc.x, lo.x, hi.x, t0.z, t0.w);
            take(sc.y, lo.y, hi.y, t1.z, t1.w);
            take(sc.z, lo.z, hi.z, t2.z, t2.w);
            take(sc.w, lo.w, hi.w, t3.z, t3.w);
        }
    } else {
        for (uint32_t k = 0; k < a.per_region; ++k) {
            const uint64_t h = h0 + k;
            take(a.xscore[h], a.xlo[h], a.xhi[h], a.hits[4 * h + 2], a.hits[4 * h + 3]);
        }
    }

    const int64_t mm = (int64_t)m, nn = (int64_t)nwin;
    const int64_t blo = J.band_lo > -(mm - 1) ? J.band_lo : -(mm - 1);
    const int64_t bhi = J.band_hi < nn - 1 ? J.band_hi : nn - 1;
    const float width = bhi >= blo ? (float)(bhi - blo + 1) : 0.f;
    const float fb = (float)best;
    const float pr = a.per_region ? (float)a.per_region : 1.f;
    float f[64];
    f[0] = fb / 64.f;
    f[1] = logf(1.f + fmaxf(fb, 0.f));
    f[2] = (float)hsp / pr;
    f[3] = (float)hsum / 128.f;
    f[4] = hsp >= 2 ? (float)(dmax - dmin) / 64.f : 0.f;
    f[5] = width / 128.f;
    f[6] = logf(1.f + (float)nwin);
    f[7] = logf(1.f + (float)m);
    f[8] = fb / logf(2.f + (float)m * (float)nwin);
    f[9] = a.per_region ? ssum / pr / 64.f : 0.f;
    f[10] = (float)ext / 64.f;
    f[11] = fb / (1.f + (float)ext);
    const float4* qc = reinterpret_cast<const float4*>(a.qcomp + (uint64_t)J.query * 20);   // 80 B rows
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const float4 v = qc[x];
        f[12 + 4 * x] = v.x;
        f[13 + 4 * x] = v.y;
        f[14 + 4 * x] = v.z;
        f[15 + 4 * x] = v.w;
    }
#pragma unroll
    for (int c = 0; c < 20; ++c) f[32 + c] = nwin ? (float)H[c] / (float)nwin : 0.f;
#pragma unroll
    for (int c = 52; c < 64; ++c) f[c] = 0.f;
    store_features(f, a.mean, a.inv_std, a.out + r * 64);
}

// Features of the search's candidate regions (saber_hip_search): the
// evidence is the region's own (best HSP, HSP count and saturating sum,
// diagonal span), with its plan's window and band, the lengths, the query's
// region threshold and the query / window compositions:
//   0 best / 64   1 ln(1 + best)   2 HSPs / 4   3 HSP sum / 128
//   4 diagonal span / 64   5 band width / 128 (the window rectangle's when
//   unbanded)   6 ln(1 + window)   7 ln(1 + query)   8 best / ln(2 + query x
//   window)   9 mean HSP score / 64   10 banded   11 best / region threshold
//   12..31 query composition   32..51 window composition   52 ln(1 + subject)
//   53 (best - region threshold) / 64   54 HSP sum / region threshold
//   55..63 zero
__global__ __launch_bounds__(256) void search_features_kernel(SearchFeatureArgs a) {
    __shared__ uint32_t hist[256 * kHistStride];
    uint32_t* H = hist + threadIdx.x * kHistStride;
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    const SearchFeatRegion R = a.regions[r];
    window_hist(a.sres + R.s_off, R.n, H);
    const float fb = (float)R.best, thr = fmaxf(R.thr, 1.f);
    const float width = R.band_hi >= R.band_lo ? (float)(R.band_hi - R.band_lo + 1) : 0.f;
    float f[64];
    f[0] = fb / 64.f;
    f[1] = logf(1.f + fmaxf(fb, 0.f));
    f[2] = (float)R.n_hsp / 4.f;
    f[3] = (float)R.hsp_sum / 128.f;
    f[4] = (float)R.d_span / 64.f;
    f[5] = width / 128.f;
    f[6] = logf(1.f + (float)R.n);
    f[7] = logf(1.f + (float)R.m);
    f[8] = fb / logf(2.f + (float)R.m * (float)R.n);
    f[9] = R.n_hsp ? (float)R.hsp_sum / (float)R.n_hsp / 64.f : 0.f;
    f[10] = R.banded ? 1.f : 0.f;
    f[11] = fb / thr;
    const float4* qc = reinterpret_cast<const float4*>(a.qcomp + (uint64_t)R.query * 20);
#pragma unroll
    for (int x = 0; x < 5; ++x) {
        const float4 v = qc[x];
        f[12 + 4 * x] = v.x;
        f[13 + 4 * x] = v.y;
        f[14 + 4 * x] = v.z;
        f[15 + 4 * x] = v.w;
    }
#pragma unroll
    for (int c = 0; c < 20; ++c) f[32 + c] = R.n ? (float)H[c] / (float)R.n : 0.f;
    f[52] = logf(1.f + (float)R.slen);
    f[53] = (fb - R.thr) / 64.f;
    f[54] = (float)R.hsp_sum / thr;
#pragma unroll
    for (int c = 55; c < 64; ++c) f[c] = 0.f;
    store_features(f, a.mean, a.inv_std, a.out + r * 64);
}

hipError_t launch_search_features(const SearchFeatureArgs& a, const uint64_t* qoff, const uint8_t* qres, uint32_t nq,
                                  hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    SABER_LAUNCH(query_composition_kernel, dim3((uint32_t)std::min<uint64_t>((nq + 3) / 4, 4096)), dim3(256),
                       0, s, qoff, qres, nq, a.qcomp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SABER_LAUNCH(search_features_kernel, dim3((uint32_t)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_region_features(const FeatureArgs& a, uint32_t nq, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    SABER_LAUNCH(query_composition_kernel, dim3((uint32_t)std::min<uint64_t>((nq + 3) / 4, 4096)), dim3(256),
                       0, s, a.qoff, a.qres, nq, a.qcomp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SABER_LAUNCH(region_features_kernel, dim3((uint32_t)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
