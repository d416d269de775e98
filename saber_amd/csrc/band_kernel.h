// Packed-int16 banded Smith-Waterman score kernel for gfx950.
//
// Semantics: exactly banded_score (src/algorithm/banded.rs:134-224, 408-449):
// the best score of a local alignment constrained to the diagonal band
// lo <= j - i <= lo + w - 1, the first row-major cell holding it (exclusive
// ends), and the number of band cells inside the rectangle.
//
// Layout of the work:
//  * Two independent jobs share each lane: every DP value is a pair of int16
//    (job A in the low half, job B in the high half) and every DP operation is
//    one v_pk_{add,sub,max,min}_i16 with clamp.
//  * The DP runs row by row (query position i outer). The band cells of a row
//    live in registers in band coordinates (cell k <-> subject column
//    j = i + lo + k), so the diagonal predecessor is the same register and
//    the vertical predecessor the next one: no shuffles inside a lane.
//  * L lanes cooperate on one job pair, each owning C consecutive band cells.
//    Lane l works on row t - l at step t (a one-row lag per lane), so the
//    only cross-lane values are the left neighbour's (H, E) from the previous
//    step and the right neighbour's (H, F) of its first cell of the current
//    step, both moved with DPP row shifts.
//  * Substitution scores come from a 32 x 32 int16 table in LDS indexed by the
//    query code (per row) and the subject code (per cell, kept as pre-doubled
//    bytes in a register window that shifts by one column per row). Column 31
//    is a sentinel for subject positions outside the window: its score
//    (-32768) makes every cell left of the rectangle exactly 0 (the border
//    value the scalar kernel uses) and keeps cells right of it from ever
//    influencing a real cell.
//  * Cells beyond the job's clipped width (k >= w) are capped to -32768 for H
//    and F; only the last K cells of a band can be beyond, so K cap registers
//    suffice.
//  * Exactness: every H/E/F of a local alignment is bounded by the optimum; the
//    only upward saturation is H_diag + S + bias > 32767, which forces that
//    cell's H to 32767. A job whose best reaches 32767 is therefore flagged and
//    rescored by the exact int32 kernel; every other job is exact.
//  * End cell: each lane keeps the best of its own cells and snapshots its C
//    cells (to private scratch) whenever that best strictly improves; at the
//    end the group takes the global best, the earliest row that reached it and
//    the leftmost cell of that row.
#pragma once
#include "common.h"

namespace saber {

typedef short v2s __attribute__((ext_vector_type(2)));

__device__ __forceinline__ v2s vadd(v2s a, v2s b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ v2s vsub(v2s a, v2s b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ v2s vmax(v2s a, v2s b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ v2s vmin(v2s a, v2s b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ uint32_t bits(v2s v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ v2s pair(uint32_t v) { return __builtin_bit_cast(v2s, v); }

// Lane i receives lane i-1 / i+1 of its row of 16 (callers override the
// values that cross a group boundary).
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, false);
}

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Raw residue at window column j, read from a clamped (always valid) address
// so the load never sits behind a branch; subject_code() turns it into the
// pre-doubled code, or the sentinel when j is outside [0, n).
__device__ __forceinline__ uint32_t subject_raw(const uint8_t* s, int64_t j, uint32_t n) {
    int64_t jc = j < 0 ? 0 : j;
    const int64_t top = n > 0 ? (int64_t)n - 1 : 0;
    jc = jc > top ? top : jc;
    return s[jc];
}
__device__ __forceinline__ uint32_t subject_code(uint32_t raw, int64_t j, uint32_t n) {
    if (j < 0 || j >= (int64_t)n) return 2u * kSentinel;
    return 2u * (raw < 24 ? raw : (uint32_t)kCodeX);
}

template <int L, int C, int K, bool BIAS>
__global__ __launch_bounds__(256) void band_pair_kernel(BandKernelArgs a) {
    static_assert(C >= 2 && K >= 1 && K <= C, "variant shape");
    static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "groups must tile a DPP row");
    constexpr int NW = (C + 3) / 4;
    __shared__ short T[kTableDim * kTableDim];
    for (int t = threadIdx.x; t < kTableDim * kTableDim; t += blockDim.x) T[t] = a.table[t];
    __syncthreads();
    const char* Tb = reinterpret_cast<const char*>(T);

    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = gl / L;
    const int l = (int)(gl % L);
    if (g >= a.npairs) return;

    uint32_t jid[2];
    int64_t qoff[2], soff[2];
    uint32_t m[2], n[2];
    int lo[2], w[2], ifirst[2], rows[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        jid[h] = a.pairs[2 * g + h];
        if (jid[h] != kNone) {
            const BandJobDev J = a.jobs[jid[h]];
            qoff[h] = (int64_t)J.q_off;
            soff[h] = (int64_t)J.s_off;
            m[h] = J.m;
            n[h] = J.n;
            lo[h] = J.lo;
            w[h] = J.w;
            ifirst[h] = J.i_first;
            rows[h] = J.rows;
        } else {
            qoff[h] = 0;
            soff[h] = 0;
            m[h] = 1;
            n[h] = 0;
            lo[h] = 0;
            w[h] = 1;
            ifirst[h] = 0;
            rows[h] = 0;
        }
    }
    const int R = rows[0] > rows[1] ? rows[0] : rows[1];
    const uint8_t* qp[2] = {a.qres + qoff[0], a.qres + qoff[1]};
    const uint8_t* sp[2] = {a.sres + soff[0], a.sres + soff[1]};
    const int8_t* bp[2] = {nullptr, nullptr};
    if constexpr (BIAS) {
        bp[0] = a.qbias + qoff[0];
        bp[1] = a.qbias + qoff[1];
    }

    const v2s GO = {(short)a.go, (short)a.go};
    const v2s GE = {(short)a.ge, (short)a.ge};
    const v2s NEG = {kNeg16, kNeg16};
    const v2s ZERO = {0, 0};

    // Caps for the last K cells of this lane: cell k stays if k < w, else -inf.
    v2s cap[K];
#pragma unroll
    for (int x = 0; x < K; ++x) {
        const int k = l * C + (C - K) + x;
        cap[x] = v2s{k < w[0] ? kMax16 : kNeg16, k < w[1] ? kMax16 : kNeg16};
    }

    v2s H[C], F[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H[c] = ZERO;
        F[c] = NEG;
    }

    // Subject window of this lane for row 0 (pre-doubled codes, 4 per word).
    uint32_t Wd[2][NW];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int x = 0; x < NW; ++x) Wd[h][x] = 0;
        const int64_t j0 = (int64_t)ifirst[h] + lo[h] + l * C;
        uint32_t raw[C];
#pragma unroll
        for (int c = 0; c < C; ++c) raw[c] = subject_raw(sp[h], j0 + c, n[h]);
#pragma unroll
        for (int c = 0; c < C; ++c) Wd[h][c >> 2] |= subject_code(raw[c], j0 + c, n[h]) << (8 * (c & 3));
    }

    // Per-lane byte streams for the row this lane prepares next (row k = r + 1
    // at step r): the subject column entering the window, the query code and
    // the bias. Each is a dword FIFO (cur, n1, n2) refilled every 4 rows with
    // an unaligned load issued 8 rows ahead, so row k's bytes were requested
    // two refills earlier. Device buffers carry 16 guard bytes in front and 64
    // behind, and every load address is clamped into them; bytes from outside
    // a sequence are replaced (sentinel column / garbage rows never used).
    auto load4 = [](const uint8_t* buf, int64_t off, uint64_t len) -> uint32_t {
        const int64_t top = (int64_t)len + 60;
        off = off < -16 ? -16 : (off > top ? top : off);
        uint32_t v;
        __builtin_memcpy(&v, buf + off, 4);
        return v;
    };
    const int64_t jcol0[2] = {(int64_t)ifirst[0] + lo[0] + l * C + (C - 1),
                              (int64_t)ifirst[1] + lo[1] + l * C + (C - 1)};
    uint32_t s_cur[2], s_n1[2], s_n2[2], q_cur[2], q_n1[2], q_n2[2];
    uint32_t b_cur[2] = {0, 0}, b_n1[2] = {0, 0}, b_n2[2] = {0, 0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t sj = soff[h] + jcol0[h], qi = qoff[h] + ifirst[h];
        s_cur[h] = load4(a.sres, sj + 1, a.sres_len);
        s_n1[h] = load4(a.sres, sj + 5, a.sres_len);
        s_n2[h] = load4(a.sres, sj + 9, a.sres_len);
        q_cur[h] = load4(a.qres, qi + 1, a.qres_len);
        q_n1[h] = load4(a.qres, qi + 5, a.qres_len);
        q_n2[h] = load4(a.qres, qi + 9, a.qres_len);
        if constexpr (BIAS) {
            const uint8_t* bb = reinterpret_cast<const uint8_t*>(a.qbias);
            b_cur[h] = load4(bb, qi + 1, a.qres_len);
            b_n1[h] = load4(bb, qi + 5, a.qres_len);
            b_n2[h] = load4(bb, qi + 9, a.qres_len);
        }
    }

    int cbest[2] = {0, 0};
    int csnap_row[2] = {-1, -1};
    uint32_t snap[2][C];
    uint64_t cells[2] = {0, 0};
    uint32_t outHl = bits(NEG), outE = bits(NEG);

    // Row data of row 0 of this lane.
    uint32_t rb[2];
    v2s bpair = ZERO;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t c = qp[h][ifirst[h] < (int)m[h] ? ifirst[h] : 0];
        rb[h] = (c < 24 ? c : (uint32_t)kCodeX) * 64u;
    }
    if constexpr (BIAS) {
        bpair = v2s{(short)bp[0][ifirst[0] < (int)m[0] ? ifirst[0] : 0],
                    (short)bp[1][ifirst[1] < (int)m[1] ? ifirst[1] : 0]};
    }

    const int nsteps = R + L - 1;
    for (int t = 0; t < nsteps; ++t) {
        const int r = t - l;
        const bool act = (r >= 0) && (r < R);
        v2s inHl = NEG, inE = NEG;
        if constexpr (L > 1) {
            inHl = pair(from_prev_lane(outHl));
            inE = pair(from_prev_lane(outE));
            if (l == 0) {
                inHl = NEG;
                inE = NEG;
            }
        }
        v2s Hl = NEG, E = NEG, rm = ZERO;
        auto cell = [&](int c, v2s hup, v2s fup, v2s el, v2s hl, bool first) {
            const uint32_t ba = (Wd[0][c >> 2] >> (8 * (c & 3))) & 0xFFu;
            const uint32_t bb = (Wd[1][c >> 2] >> (8 * (c & 3))) & 0xFFu;
            const short sa = *reinterpret_cast<const short*>(Tb + rb[0] + ba);
            const short sb = *reinterpret_cast<const short*>(Tb + rb[1] + bb);
            v2s P = v2s{sa, sb};
            if constexpr (BIAS) P = vadd(P, bpair);
            const v2s D = vadd(H[c], P);
            v2s Fn = vmax(vsub(hup, GO), vsub(fup, GE));
            v2s En = first ? vmax(vsub(hl, GO), vsub(el, GE)) : vmax(vsub(Hl, GO), vsub(E, GE));
            v2s hv = vmax(vmax(D, En), vmax(Fn, ZERO));
            if (c >= C - K) {
                hv = vmin(hv, cap[c - (C - K)]);
                Fn = vmin(Fn, cap[c - (C - K)]);
            }
            H[c] = hv;
            F[c] = Fn;
            Hl = hv;
            E = En;
            rm = vmax(rm, hv);
        };

        if (act) {
            // Cell 0: its left neighbour is the previous lane's last cell.
            if constexpr (C > 1) cell(0, H[1], F[1], inE, inHl, true);
        }
        v2s upH = NEG, upF = NEG;
        if constexpr (L > 1) {
            upH = pair(from_next_lane(bits(H[0])));
            upF = pair(from_next_lane(bits(F[0])));
            if (l == L - 1) {
                upH = NEG;
                upF = NEG;
            }
        }
        if (act) {
#pragma unroll
            for (int c = 1; c < C - 1; ++c) cell(c, H[c + 1], F[c + 1], NEG, NEG, false);
            cell(C - 1, upH, upF, NEG, NEG, false);
            outHl = bits(Hl);
            outE = bits(E);

#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (r < rows[h]) {
                    const int i = ifirst[h] + r;
                    const int rmh = h == 0 ? (int)rm.x : (int)rm.y;
                    if (rmh > cbest[h]) {
                        cbest[h] = rmh;
                        csnap_row[h] = i;
#pragma unroll
                        for (int c = 0; c < C; ++c) snap[h][c] = bits(H[c]);
                    }
                    if (l == 0) {
                        int kmin = -(i + lo[h]);
                        if (kmin < 0) kmin = 0;
                        int kmax = (int)n[h] - 1 - i - lo[h];
                        if (kmax > w[h] - 1) kmax = w[h] - 1;
                        cells[h] += (uint64_t)(kmax - kmin + 1);
                    }
                }
            }
            // Next row's data from the streams; shift the windows one column
            // and append the new column.
            const int k = r + 1;
            short bn[2] = {0, 0};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t sraw = s_cur[h] & 0xFFu, qc = q_cur[h] & 0xFFu;
                if constexpr (BIAS) bn[h] = (short)(int8_t)(b_cur[h] & 0xFFu);
                s_cur[h] >>= 8;
                q_cur[h] >>= 8;
                if constexpr (BIAS) b_cur[h] >>= 8;
#pragma unroll
                for (int x = 0; x < NW - 1; ++x) Wd[h][x] = __builtin_amdgcn_alignbyte(Wd[h][x + 1], Wd[h][x], 1);
                constexpr int b = (C - 1) & 3;
                uint32_t last = Wd[h][NW - 1] >> 8;
                if constexpr (b > 0) last &= (1u << (8 * b)) - 1u;
                Wd[h][NW - 1] = last | (subject_code(sraw, jcol0[h] + k, n[h]) << (8 * b));
                rb[h] = (qc < 24 ? qc : (uint32_t)kCodeX) * 64u;
            }
            if constexpr (BIAS) bpair = v2s{bn[0], bn[1]};
            if ((k & 3) == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    s_cur[h] = s_n1[h];
                    s_n1[h] = s_n2[h];
                    s_n2[h] = load4(a.sres, soff[h] + jcol0[h] + k + 9, a.sres_len);
                    q_cur[h] = q_n1[h];
                    q_n1[h] = q_n2[h];
                    q_n2[h] = load4(a.qres, qoff[h] + ifirst[h] + k + 9, a.qres_len);
                    if constexpr (BIAS) {
                        b_cur[h] = b_n1[h];
                        b_n1[h] = b_n2[h];
                        b_n2[h] = load4(reinterpret_cast<const uint8_t*>(a.qbias),
                                        qoff[h] + ifirst[h] + k + 9, a.qres_len);
                    }
                }
            }
        }
    }

    // Group reduction: global best, earliest row reaching it, leftmost cell.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int best = cbest[h];
#pragma unroll
        for (int d = 1; d < L; d <<= 1) {
            const int o = __shfl_xor(best, d, L);
            best = o > best ? o : best;
        }
        uint32_t key = 0xFFFFFFFFu;
        if (best > 0 && best < kMax16 && cbest[h] == best) {
            for (int c = 0; c < C; ++c) {
                const uint32_t v = snap[h][c];
                const int hv = h == 0 ? (int)(short)(v & 0xFFFFu) : (int)(short)(v >> 16);
                if (hv == best) {
                    key = (uint32_t)csnap_row[h] * (uint32_t)(L * C) + (uint32_t)(l * C + c);
                    break;
                }
            }
        }
#pragma unroll
        for (int d = 1; d < L; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)key, d, L);
            key = o < key ? o : key;
        }
        if (l == 0 && jid[h] != kNone) {
            const uint32_t id = jid[h];
            if (best >= kMax16) {
                const uint32_t slot = atomicAdd(a.nflagged, 1u);
                a.flagged[slot] = id;
            } else {
                uint32_t qe = 0, se = 0;
                if (best > 0) {
                    const uint32_t row = key / (uint32_t)(L * C);
                    const uint32_t k = key % (uint32_t)(L * C);
                    qe = row + 1;
                    se = (uint32_t)((int)row + lo[h] + (int)k + 1);
                }
                a.out.score[id] = best;
                a.out.q_end[id] = qe;
                a.out.s_end[id] = se;
                a.out.cells[id] = cells[h];
            }
        }
    }
}

template <int L, int C, int K, bool BIAS>
hipError_t launch_band(const BandKernelArgs& a, hipStream_t s) {
    if (a.npairs == 0) return hipSuccess;
    const uint64_t threads = (uint64_t)a.npairs * L;
    const uint32_t block = 256;
    const uint32_t grid = (uint32_t)((threads + block - 1) / block);
    hipLaunchKernelGGL((band_pair_kernel<L, C, K, BIAS>), dim3(grid), dim3(block), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
