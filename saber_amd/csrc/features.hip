// Feature vectors of candidate regions for the learned prefilter, computed on
// the device from what the pipeline holds when the prefilter runs (after the
// X-drop stage and the window/band plan, before banded scoring): the region's
// X-drop evidence, its geometry and the residue composition of the query and
// the subject window. One wave per region.
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

__global__ __launch_bounds__(256) void region_features_kernel(FeatureArgs a) {
    __shared__ uint32_t hist[4][40];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < a.n; r += nw) {
        const saber_hip_band_job J = a.jobs[r];
        if (lane < 40) hist[wv][lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint8_t* q = a.qres + a.qoff[J.query];
        const uint64_t m = a.qoff[J.query + 1] - a.qoff[J.query];
        const uint8_t* s = a.sres + a.soff[J.subject] + J.win_lo;
        const uint64_t nwin = J.win_hi - J.win_lo;
        for (uint64_t i = lane; i < m; i += 64) {
            const uint32_t c = q[i];
            if (c < 20) atomicAdd(&hist[wv][c], 1u);
        }
        for (uint64_t i = lane; i < nwin; i += 64) {
            const uint32_t c = s[i];
            if (c < 20) atomicAdd(&hist[wv][20 + c], 1u);
        }
        // X-drop evidence of the region's hits (lane k: hit k)
        int32_t sc = -2147483647, hsp = 0, hsum = 0, ext = 0;
        int64_t dmin = INT64_MAX, dmax = INT64_MIN;
        float ssum = 0.f;
        if ((uint32_t)lane < a.per_region) {
            const uint64_t h = r * a.per_region + lane;
            sc = a.xscore[h];
            ssum = (float)sc;
            ext = (int32_t)(a.xhi[h] - a.xlo[h]);
            if (sc >= a.hsp_threshold) {
                hsp = 1;
                hsum = sc;
                const int64_t d = (int64_t)a.hits[4 * h + 3] - (int64_t)a.hits[4 * h + 2];
                dmin = d;
                dmax = d;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            sc = max(sc, __shfl_xor(sc, o));
            hsp += __shfl_xor(hsp, o);
            hsum += __shfl_xor(hsum, o);
            ext = max(ext, __shfl_xor(ext, o));
            ssum += __shfl_xor(ssum, o);
            dmin = min(dmin, (int64_t)__shfl_xor((long long)dmin, o));
            dmax = max(dmax, (int64_t)__shfl_xor((long long)dmax, o));
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int64_t mm = (int64_t)m, nn = (int64_t)nwin;
        const int64_t blo = J.band_lo > -(mm - 1) ? J.band_lo : -(mm - 1);
        const int64_t bhi = J.band_hi < nn - 1 ? J.band_hi : nn - 1;
        const float width = bhi >= blo ? (float)(bhi - blo + 1) : 0.f;
        const float best = a.per_region ? (float)sc : 0.f;
        const float pr = a.per_region ? (float)a.per_region : 1.f;
        float f = 0.f;
        switch (lane) {
            case 0: f = best / 64.f; break;
            case 1: f = logf(1.f + fmaxf(best, 0.f)); break;
            case 2: f = (float)hsp / pr; break;
            case 3: f = (float)hsum / 128.f; break;
            case 4: f = hsp >= 2 ? (float)(dmax - dmin) / 64.f : 0.f; break;
            case 5: f = width / 128.f; break;
            case 6: f = logf(1.f + (float)nwin); break;
            case 7: f = logf(1.f + (float)m); break;
            case 8: f = best / logf(2.f + (float)m * (float)nwin); break;
            case 9: f = a.per_region ? ssum / pr / 64.f : 0.f; break;
            case 10: f = (float)ext / 64.f; break;
            case 11: f = best / (1.f + (float)ext); break;
            default:
                if (lane >= 12 && lane < 32)
                    f = m ? (float)hist[wv][lane - 12] / (float)m : 0.f;
                else if (lane >= 32 && lane < 52)
                    f = nwin ? (float)hist[wv][lane - 12] / (float)nwin : 0.f;
        }
        if (a.mean) f = (f - a.mean[lane]) * a.inv_std[lane];
        a.out[r * 64 + lane] = (__bf16)f;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

hipError_t launch_region_features(const FeatureArgs& a, int grid, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(region_features_kernel, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
