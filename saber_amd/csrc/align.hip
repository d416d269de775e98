// Banded Smith-Waterman with traceback on the device: banded_align_until
// (src/algorithm/banded.rs:472-521), i.e. the scalar recurrence of
// banded_align_scalar (banded.rs:227-333) with its direction bits
// (org | e_ext << 2 | f_ext << 3: origin 0 start, 1 diagonal, 2 E, 3 F; ties
// diagonal > E > F, extension only when strictly better), the row stop of the
// AVX2 kernels the dispatcher runs (banded_simd.rs:324, 594: after the first
// row holding band cells whose running best reaches stop_at), and walk
// (banded.rs:336-403).
//
// One thread per job, exact int32: the two band rows in LDS or a per-job
// slice of a work buffer (below), the direction bytes of the job's band rows
// (rows i_first .. i_first + rows - 1, w bytes each) in a per-job slice of a
// direction buffer.
// The walk writes the reversed operations into a per-job byte slice of m + n
// bytes, then the run-length operations in alignment order at the same
// offset of the output arrays.
#include "common.h"

namespace saber {

constexpr int32_t kNegAlign = INT32_MIN / 4;

__device__ __forceinline__ int32_t amax(int32_t a, int32_t b) { return a > b ? a : b; }

// Band rows updated in place: cell k of the current row reads the previous
// row's H[k] (diagonal) and H[k+1], F[k+1] (from above) before it writes H[k],
// F[k]; cells below the row's first band cell are never valid (the first
// cell only moves left from row to row), cells past its last are reset after
// the row. Jobs of band width <= kLdsWidth keep the two rows in LDS (2 KB a
// thread), wider ones in a per-job slice of a work buffer. A group's subject
// codes are one dword load (the residue buffers carry guard bytes past their
// end), issued one group ahead.
constexpr int kLdsWidth = kAlignLdsWidth;
constexpr int kLdsThreads = 64;

template <bool LDS>
__global__ __launch_bounds__(kLdsThreads) void band_align_kernel(AlignArgs a) {
    __shared__ int32_t lrows[LDS ? kLdsThreads * 2 * (kLdsWidth + 1) : 1];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t id = a.ids[t];
    const BandJobDev J = a.jobs[id];
    AlignRecord rec{};
    rec.status = 1;
    if (J.w <= 0 || J.rows <= 0) {
        a.out[id] = rec;
        return;
    }
    const int64_t w = J.w, lo = J.lo, n = J.n;
    int32_t* H = LDS ? lrows + threadIdx.x * 2 * (kLdsWidth + 1) : a.work + a.work_off[t];
    int32_t* F = H + (LDS ? (kLdsWidth + 1) : (w + 1));
    for (int64_t k = 0; k <= w; ++k) {
        H[k] = kNegAlign;
        F[k] = kNegAlign;
    }
    uint8_t* dir = a.dir + a.dir_off[t];
    const uint8_t* q = a.qres + J.q_off;
    const uint8_t* s = a.sres + J.s_off;
    const int8_t* bias = a.qbias ? a.qbias + J.q_off : nullptr;
    const int32_t go = a.go, ge = a.ge;
    const int32_t stop = a.stop_at ? a.stop_at[id] : INT32_MAX;
    int32_t best = 0;
    int64_t bi = 0, bk = 0;
    uint64_t cells = 0;
    int64_t kmax_prev = w - 1;
    const int64_t i_end = (int64_t)J.i_first + J.rows;
    for (int64_t i = J.i_first; i < i_end; ++i) {
        int64_t kmin = -(i + lo);
        if (kmin < 0) kmin = 0;
        int64_t kmax = n - 1 - i - lo;
        if (kmax > w - 1) kmax = w - 1;
        const uint32_t qc = q[i] < 24 ? q[i] : (uint32_t)kCodeX;
        const int32_t* row = a.mat + qc * 24;
        const int32_t b = bias ? (int32_t)bias[i] : 0;
        uint8_t* drow = dir + (i - J.i_first) * w;
        int32_t e = kNegAlign, hl = kNegAlign;
        int32_t hnext = H[kmin];   // previous row's H at the next cell
        uint32_t snext;            // subject codes of the next group
        __builtin_memcpy(&snext, s + (i + lo + kmin), 4);
        for (int64_t k0 = kmin; k0 <= kmax; k0 += 4) {
            const int nk = (int)(kmax - k0 + 1 < 4 ? kmax - k0 + 1 : 4);
            int32_t hpv[5], fpv[4];
            uint32_t scv[4];
            const uint32_t sw = snext;
            if (k0 + 4 <= kmax) __builtin_memcpy(&snext, s + (i + lo + k0 + 4), 4);
            hpv[0] = hnext;
#pragma unroll
            for (int u = 1; u < 5; ++u) hpv[u] = H[k0 + (u < nk ? u : nk)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + (u < nk ? u : nk - 1);
                fpv[u] = F[k + 1];
                const uint32_t sv = (sw >> (8 * u)) & 0xFFu;
                scv[u] = sv < 24 ? sv : (uint32_t)kCodeX;
            }
            hnext = hpv[nk];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (u >= nk) break;
                const int64_t k = k0 + u;
                const int64_t j = i + lo + k;
                const int32_t eo = hl - go, ee = e - ge;
                const uint32_t e_ext = ee > eo;
                e = k > kmin ? amax(ee, eo) : kNegAlign;
                const int32_t fo = hpv[u + 1] - go, fe = fpv[u] - ge;
                const uint32_t f_ext = fe > fo;
                const int32_t f = i >= 1 ? amax(fe, fo) : kNegAlign;
                const int32_t hd = (i >= 1 && j >= 1) ? hpv[u] : 0;
                const int32_t diag = hd + row[scv[u]] + b;
                const uint32_t c1 = diag > 0;
                const int32_t h1 = amax(diag, 0);
                const uint32_t c2 = e > h1;
                const int32_t h2 = amax(h1, e);
                const uint32_t c3 = f > h2;
                const int32_t h = amax(h2, f);
                const uint32_t org = c3 ? 3u : (c2 ? 2u : c1);
                H[k] = h;
                F[k] = f;
                drow[k] = (uint8_t)(org | (e_ext << 2) | (f_ext << 3));
                hl = h;
                if (h > best) {
                    best = h;
                    bi = i;
                    bk = k;
                }
            }
        }
        // the previous row's cells past this row's last band cell
        for (int64_t k = kmax + 1; k <= kmax_prev; ++k) {
            H[k] = kNegAlign;
            F[k] = kNegAlign;
        }
        kmax_prev = kmax;
        cells += (uint64_t)(kmax - kmin + 1);
        if (best >= stop) break;
    }
    rec.cells = cells;
    if (best <= 0) {
        a.out[id] = rec;
        return;
    }
    // walk: operations from the end cell backwards
    uint8_t* rev = a.op_raw + a.op_off[t];
    int64_t nrev = 0;
    int64_t i = bi, k = bk;
    int state = 0;   // 0 = H, 1 = E (query gap, 'D'), 2 = F (subject gap, 'I')
    for (;;) {
        const uint8_t d = dir[(i - J.i_first) * w + k];
        if (state == 0) {
            const uint32_t org = d & 3u;
            if (org == 0) break;
            if (org == 1) {
                rev[nrev++] = 'M';
                if (i == 0 || i + lo + k == 0) break;
                i -= 1;
            } else {
                state = org == 2 ? 1 : 2;
            }
        } else if (state == 1) {
            rev[nrev++] = 'D';
            k -= 1;
            if (!(d & 4u)) state = 0;
        } else {
            rev[nrev++] = 'I';
            i -= 1;
            k += 1;
            if (!(d & 8u)) state = 0;
        }
    }
    const int64_t qe = bi + 1, se = bi + lo + bk + 1;
    int64_t qs = qe, ss = se;
    uint8_t* oc = a.op_codes + a.op_off[t];
    uint32_t* ol = a.op_lens + a.op_off[t];
    uint32_t nr = 0;
    for (int64_t x = nrev - 1; x >= 0; --x) {
        const uint8_t op = rev[x];
        if (op == 'M') {
            --qs;
            --ss;
        } else if (op == 'I') {
            --qs;
        } else {
            --ss;
        }
        if (nr > 0 && oc[nr - 1] == op) {
            ++ol[nr - 1];
        } else {
            oc[nr] = op;
            ol[nr] = 1;
            ++nr;
        }
    }
    rec.status = 0;
    rec.score = best;
    rec.q_start = (uint32_t)qs;
    rec.q_end = (uint32_t)qe;
    rec.s_start = (uint32_t)ss;
    rec.s_end = (uint32_t)se;
    rec.nops = nr;
    a.out[id] = rec;
}

hipError_t launch_band_align(const AlignArgs& a, bool lds, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint32_t grid = (a.n + kLdsThreads - 1) / kLdsThreads;
    if (lds)
        hipLaunchKernelGGL(band_align_kernel<true>, dim3(grid), dim3(kLdsThreads), 0, s, a);
    else
        hipLaunchKernelGGL(band_align_kernel<false>, dim3(grid), dim3(kLdsThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
