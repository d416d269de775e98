// Packed-int16 full-matrix Smith-Waterman kernel for gfx950 (exhaustive mode).
//
// Semantics: score_many's score (src/algorithm/sw_simd.rs:95-153) with the end
// cell of sw_reference (src/algorithm/reference.rs:15-52): the first row-major
// optimal cell (exclusive ends), plus the m*n cell count.
//
// Work: the cross product of a query set and a subject set. Subjects are
// bucketed by length into variants; two subjects of similar length share each
// lane (int16 halves), both against the same query (blockIdx.y), so the query
// residue of a row is common to the pair.
//
// Layout: rows are query positions, columns subject positions. L lanes serve a
// subject pair, lane l owning columns [l*C, l*C + C) and working on row t - l
// at step t. A cell needs its left neighbour (same row; for the first column of
// a lane, the previous lane's last cell from the previous step), its upper
// neighbour (own register from the previous row) and its diagonal (the upper
// neighbour of the left cell: the previous lane's last cell from two steps
// back), so only the left lane's (H, E) cross lanes, once per step, by DPP.
// Subject codes are fixed per lane and held in registers for the whole job;
// columns past the subject's end read the sentinel column (-32768), which keeps
// them from influencing any real cell. Saturation and the end cell are handled
// as in the band kernel.
#pragma once
#include "band_kernel.h"

namespace saber {

// Raw residue at subject column j, read from a clamped (always valid) address
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

// Lane i receives lane i-1: within a 16-lane row for groups that tile rows,
// across the whole wave (wave_shr:1) for 32- and 64-lane groups.
template <int L>
__device__ __forceinline__ uint32_t full_from_prev_lane(uint32_t v) {
    if constexpr (L <= 16)
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    else
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);
}

template <int L, int C, bool BIAS>
__global__ __launch_bounds__(256) void full_pair_kernel(FullKernelArgs a) {
    static_assert(C >= 2, "variant shape");
    static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64,
                  "groups must tile a wave");
    constexpr int NW = (C + 3) / 4;
    __shared__ short T[kTableDim * kTableDim];
    for (int t = threadIdx.x; t < kTableDim * kTableDim; t += blockDim.x) T[t] = a.table[t];
    __syncthreads();
    const char* Tb = reinterpret_cast<const char*>(T);

    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = gl / L;
    const int l = (int)(gl % L);
    if (g >= a.npairs) return;
    const uint32_t slot = a.slot0 + blockIdx.y;
    const uint32_t qid = a.queries[slot];
    const int64_t q0 = (int64_t)a.qoff[qid];
    const int m = (int)(a.qoff[qid + 1] - a.qoff[qid]);

    uint32_t sid[2];
    uint32_t n[2];
    const uint8_t* sp[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        sid[h] = a.pairs[2 * g + h];
        if (sid[h] != kNone) {
            const uint64_t s0 = a.soff[sid[h]];
            n[h] = (uint32_t)(a.soff[sid[h] + 1] - s0);
            sp[h] = a.sres + s0;
        } else {
            n[h] = 0;
            sp[h] = a.sres;
        }
    }
    const int R = m;  // every row of the query holds cells of both jobs

    const v2s GO = {(short)a.go, (short)a.go};
    const v2s GE = {(short)a.ge, (short)a.ge};
    const v2s NEG = {kNeg16, kNeg16};
    const v2s ZERO = {0, 0};

    v2s H[C], F[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H[c] = ZERO;
        F[c] = NEG;
    }
    uint32_t Wd[2][NW];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int x = 0; x < NW; ++x) Wd[h][x] = 0;
        const int64_t j0 = (int64_t)l * C;
        uint32_t raw[C];
#pragma unroll
        for (int c = 0; c < C; ++c) raw[c] = subject_raw(sp[h], j0 + c, n[h]);
#pragma unroll
        for (int c = 0; c < C; ++c) Wd[h][c >> 2] |= subject_code(raw[c], j0 + c, n[h]) << (8 * (c & 3));
    }

    auto load4 = [](const uint8_t* buf, int64_t off, uint64_t len) -> uint32_t {
        const int64_t top = (int64_t)len + 60;
        off = off < -16 ? -16 : (off > top ? top : off);
        uint32_t v;
        __builtin_memcpy(&v, buf + off, 4);
        return v;
    };
    // Query code / bias streams for rows 1, 2, ... (row 0 read directly).
    uint32_t q_cur = load4(a.qres, q0 + 1, a.qres_len), q_n1 = load4(a.qres, q0 + 5, a.qres_len),
             q_n2 = load4(a.qres, q0 + 9, a.qres_len);
    uint32_t b_cur = 0, b_n1 = 0, b_n2 = 0;
    const uint8_t* bb = reinterpret_cast<const uint8_t*>(a.qbias);
    if constexpr (BIAS) {
        b_cur = load4(bb, q0 + 1, a.qres_len);
        b_n1 = load4(bb, q0 + 5, a.qres_len);
        b_n2 = load4(bb, q0 + 9, a.qres_len);
    }
    uint32_t qc0 = m > 0 ? a.qres[q0] : 0;
    uint32_t rb = (qc0 < 24 ? qc0 : (uint32_t)kCodeX) * 64u;
    v2s bpair = ZERO;
    if constexpr (BIAS) {
        const short b0 = m > 0 ? (short)a.qbias[q0] : (short)0;
        bpair = v2s{b0, b0};
    }

    int cbest[2] = {0, 0};
    int csnap_row[2] = {-1, -1};
    uint32_t snap[2][C];
    uint32_t outHl = 0u, outE = bits(NEG);   // lane 0's left border: H = 0, E = -inf
    uint32_t diagIn = 0u;                    // H(i-1, l*C-1) for the current row

    const int nsteps = R + L - 1;
    for (int t = 0; t < nsteps; ++t) {
        const int r = t - l;
        const bool act = (r >= 0) && (r < R);
        v2s inHl = ZERO, inE = NEG;
        if constexpr (L > 1) {
            inHl = pair(full_from_prev_lane<L>(outHl));
            inE = pair(full_from_prev_lane<L>(outE));
            if (l == 0) {
                inHl = ZERO;
                inE = NEG;
            }
        }
        if (act) {
            v2s Hl = inHl, E = inE, rm = ZERO;
            v2s hd = pair(diagIn);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const uint32_t ba = (Wd[0][c >> 2] >> (8 * (c & 3))) & 0xFFu;
                const uint32_t bq = (Wd[1][c >> 2] >> (8 * (c & 3))) & 0xFFu;
                const short sa = *reinterpret_cast<const short*>(Tb + rb + ba);
                const short sb = *reinterpret_cast<const short*>(Tb + rb + bq);
                v2s P = v2s{sa, sb};
                if constexpr (BIAS) P = vadd(P, bpair);
                const v2s up = H[c];
                const v2s D = vadd(hd, P);
                const v2s Fn = vmax(vsub(up, GO), vsub(F[c], GE));
                const v2s En = vmax(vsub(Hl, GO), vsub(E, GE));
                const v2s hv = vmax(vmax(D, En), vmax(Fn, ZERO));
                hd = up;
                H[c] = hv;
                F[c] = Fn;
                Hl = hv;
                E = En;
                rm = vmax(rm, hv);
            }
            outHl = bits(Hl);
            outE = bits(E);
            diagIn = bits(inHl);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (r < m && sid[h] != kNone) {
                    const int rmh = h == 0 ? (int)rm.x : (int)rm.y;
                    if (rmh > cbest[h]) {
                        cbest[h] = rmh;
                        csnap_row[h] = r;
#pragma unroll
                        for (int c = 0; c < C; ++c) snap[h][c] = bits(H[c]);
                    }
                }
            }
            const int k = r + 1;
            const uint32_t qc = q_cur & 0xFFu;
            rb = (qc < 24 ? qc : (uint32_t)kCodeX) * 64u;
            q_cur >>= 8;
            if constexpr (BIAS) {
                const short bn = (short)(int8_t)(b_cur & 0xFFu);
                bpair = v2s{bn, bn};
                b_cur >>= 8;
            }
            if ((k & 3) == 0) {
                q_cur = q_n1;
                q_n1 = q_n2;
                q_n2 = load4(a.qres, q0 + k + 9, a.qres_len);
                if constexpr (BIAS) {
                    b_cur = b_n1;
                    b_n1 = b_n2;
                    b_n2 = load4(bb, q0 + k + 9, a.qres_len);
                }
            }
        } else if (r < 0) {
            // Keep the left lane's previous-step H: it is the diagonal of
            // this lane's first row when that row starts.
            diagIn = l == 0 ? 0u : bits(inHl);
        }
    }

    // Group reduction: global best, earliest row reaching it, leftmost column.
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
        if (l == 0 && sid[h] != kNone) {
            const uint64_t o = (uint64_t)slot * a.nsub + sid[h];
            if (best >= kMax16) {
                const uint32_t f = atomicAdd(a.nflagged, 1u);
                a.flagged[2 * f] = slot;
                a.flagged[2 * f + 1] = sid[h];
            } else {
                uint32_t qe = 0, se = 0;
                if (best > 0) {
                    qe = key / (uint32_t)(L * C) + 1;
                    se = key % (uint32_t)(L * C) + 1;
                }
                a.score[o] = best;
                if (a.q_end) a.q_end[o] = qe;
                if (a.s_end) a.s_end[o] = se;
            }
        }
    }
}

template <int L, int C, bool BIAS>
hipError_t launch_full(const FullKernelArgs& a, uint32_t nslots, hipStream_t s) {
    if (a.npairs == 0 || nslots == 0) return hipSuccess;
    const uint64_t threads = (uint64_t)a.npairs * L;
    const uint32_t block = 256;
    const uint32_t grid = (uint32_t)((threads + block - 1) / block);
    hipLaunchKernelGGL((full_pair_kernel<L, C, BIAS>), dim3(grid, nslots), dim3(block), 0, s, a);
    return hipGetLastError();
}

}  // namespace saber
