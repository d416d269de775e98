// Packed-int16 full-matrix Smith-Waterman kernel for gfx950 (exhaustive mode).
//
// Semantics: score_many's score (src/algorithm/sw_simd.rs:95-153), and with
// ENDS the end cell of sw_reference (src/algorithm/reference.rs:15-52): the
// first row-major optimal cell (exclusive ends).
//
// Work: the cross product of a query set and a subject set. Subjects are
// bucketed by length into shapes; two subjects of similar length share each
// lane (int16 halves), both against the same query, so the query residue of a
// row is common to the pair. One persistent kernel per register class drains
// a chunk list covering every (shape, query) segment.
//
// Layout: rows are query positions, columns subject positions. L lanes serve a
// subject pair, lane l owning columns [l*C, l*C + C) and working on row t - l
// at step t. A cell needs its left neighbour (same row; for the first column of
// a lane, the previous lane's last cell from the previous step), its upper
// neighbour (own register from the previous row) and its diagonal (the upper
// neighbour of the left cell: the previous lane's last cell from two steps
// back), so only the left lane's H and E cross lanes, once per step, by DPP.
//
// Scores: both halves of a lane score against the same query residue q, so
// one LDS read of the pair table TP[q][a][b] = S(q,a) | S(q,b) << 16 gives
// both. The subject codes of a lane's columns are fixed for the whole job, so
// each column's entry offset (a*100 + b*4, 16 bits) is computed once; a cell's
// lookup is one add to the row's base and one LDS read. Columns past a
// subject's end use code 24, whose score is -128 (the sentinel): such cells
// lie right of every real cell of their row and only receive values from real
// cells, lowered, so they can neither exceed a real cell nor come first.
//
// Recurrence with clamped gap states E' = max(E, 0), F' = max(F, 0): with
// H >= 0, H = max(0, D, E, F) = max(D, E', F'), and every "x - gap" of a
// clamped state is one unsigned saturating subtract; H - open feeds both the
// next column's E and the next row's F. Additions saturate at 32767: a job
// whose best reaches it is flagged for the exact int32 kernel.
#pragma once
#include <atomic>

#include "band_kernel.h"

namespace saber {

// Lane i receives lane i-1: within a 16-lane row for groups that tile rows,
// across the whole wave (wave_shr:1) for 32- and 64-lane groups.
template <int L>
__device__ __forceinline__ uint32_t full_from_prev_lane(uint32_t v) {
    if constexpr (L <= 16)
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    else
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);
}

// Compiled shapes (L lanes per subject pair, C subject columns per lane:
// subjects up to L*C long), by flavour; a shape's id is its position in the
// list. The scores-only flavour adds wider lanes (C = 40, 48) for every L;
// the end-cell flavour, whose snapshots cost registers, keeps C <= 32 except
// for the longest subjects.
#define SABER_FULL_SCORE_SHAPES(X)                                                              \
    X(1, 16) X(1, 24) X(1, 32) X(1, 40) X(1, 48) X(2, 32) X(2, 40) X(2, 48) X(4, 32) X(4, 40)   \
    X(4, 48) X(8, 32) X(8, 40) X(8, 48) X(16, 32) X(16, 40) X(16, 48) X(32, 32) X(32, 40)       \
    X(32, 48) X(64, 32) X(64, 40) X(64, 48)
#define SABER_FULL_END_SHAPES(X)                                                                \
    X(1, 16) X(1, 24) X(1, 32) X(2, 24) X(2, 32) X(4, 24) X(4, 32) X(8, 24) X(8, 32) X(16, 24) \
    X(16, 32) X(32, 24) X(32, 32) X(64, 24) X(64, 32) X(64, 40) X(64, 48)

// Workgroup size of a shape's register class: 12 waves (3 per SIMD: 168
// registers) for lanes of up to 32 columns, 8 waves (256 registers) for the
// wider lanes; the end-cell build keeps a row and a column per half instead
// of a copy of the lane's cells, so it needs no more registers than the
// scores-only build.
template <int C, bool ENDS>
constexpr int kFullThreads = C <= 32 ? 768 : 512;

// One subject pair (pair index g of a.pairs, lane l of its L) against query
// qid: the whole matrix, outputs at row qid of the [query][subject] arrays.
template <int L, int C, bool BIAS, bool ENDS>
__device__ __attribute__((noinline)) void full_group(const FullKernelArgs& a, lds_bytes Tpb, uint32_t qid,
                                                     uint32_t g, int l) {
    static_assert(C >= 2 && C % 2 == 0, "variant shape");
    static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64,
                  "groups must tile a wave");
    constexpr int NE = C / 2;
    const int64_t q0 = (int64_t)a.qoff[qid];
    const int m = (int)(a.qoff[qid + 1] - a.qoff[qid]);
    uint32_t sid[2], n[2];
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

    // Pair-table entry offsets of the lane's columns, two per dword.
    uint32_t E2[NE];
    {
        const int64_t j0 = (int64_t)l * C;
        uint32_t code[2][C];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int64_t j = j0 + c;
                // clamped address: the load never sits behind a branch
                const int64_t jc = j < (int64_t)n[h] ? j : (n[h] > 0 ? (int64_t)n[h] - 1 : 0);
                const uint32_t raw = sp[h][jc];
                code[h][c] = j < (int64_t)n[h] ? raw : (uint32_t)kBandSentinel;
            }
#pragma unroll
        for (int x = 0; x < NE; ++x)
            E2[x] = (code[0][2 * x] * 100u + code[1][2 * x] * 4u) |
                    ((code[0][2 * x + 1] * 100u + code[1][2 * x + 1] * 4u) << 16);
    }

    const v2s GO = {(short)a.go, (short)a.go};
    const v2s GE = {(short)a.ge, (short)a.ge};
    const v2s ZERO = {0, 0};

    v2s H[C], F[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H[c] = ZERO;
        F[c] = ZERO;
    }

    auto load4 = [](const uint8_t* buf, int64_t off, uint64_t len) -> uint32_t {
        const int64_t top = (int64_t)len + 60;
        off = off < -16 ? -16 : (off > top ? top : off);
        return gload_u32(buf + off);
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
    uint32_t rb = (m > 0 ? (uint32_t)a.qres[q0] : 0u) * kPairRowBytes;
    v2s bpair = ZERO;
    if constexpr (BIAS) {
        const short b0 = m > 0 ? (short)a.qbias[q0] : (short)0;
        bpair = v2s{b0, b0};
    }

    v2s best = ZERO;
    int brow[2] = {-1, -1};   // ENDS: row where the half's best last rose
    int bcol[2] = {C, C};     // ENDS: and the leftmost of the lane's cells holding it in that row
    uint32_t outH = 0u, outE = 0u;   // lane 0's left border: H = 0, E' = 0
    uint32_t diagIn = 0u;            // H(i-1, l*C-1) for the current row

    const int nsteps = m + L - 1;
    for (int t = 0; t < nsteps; ++t) {
        const int r = t - l;
        v2s inH = ZERO, inE = ZERO;
        if constexpr (L > 1) {
            inH = pair(full_from_prev_lane<L>(outH));
            inE = pair(full_from_prev_lane<L>(outE));
            if (l == 0) {
                inH = ZERO;
                inE = ZERO;
            }
        }
        if (r >= 0 && r < m) {
            v2s E = inE, rm = ZERO;
            const lds_bytes row = Tpb + rb;
            auto score = [&](int c) {
                const uint32_t e = (E2[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
                v2s P = __builtin_bit_cast(v2s, lds_u32(row + e));
                if constexpr (BIAS) P = vadd(P, bpair);
                return P;
            };
            // The diagonal term of column c + 1 is formed from the old H[c]
            // before column c's new H is: the old value's last use comes
            // first, so the new one can take its register (no moves to
            // restore the registers at the end of the row).
            v2s Dn = vadd(pair(diagIn), score(0));
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const v2s D = Dn;
                if (c + 1 < C) Dn = vadd(H[c], score(c + 1));
                const v2s hv = vmax(vmax(D, F[c]), E);
                const v2s hg = vsubz(hv, GO);
                E = vmax(hg, vsubz(E, GE));
                F[c] = vmax(hg, vsubz(F[c], GE));
                H[c] = hv;
                rm = vmax(rm, hv);
            }
            outH = bits(H[C - 1]);
            outE = bits(E);
            diagIn = bits(inH);
            if constexpr (ENDS) {
                // halves whose best rose: remember the row and the leftmost
                // cell of the lane holding the new best (rare after the first
                // rows, so behind a branch)
                const bool r0 = rm.x > best.x, r1 = rm.y > best.y;
                if (r0 || r1) {
                    int c0 = C, c1 = C;
#pragma unroll
                    for (int c = C - 1; c >= 0; --c) {
                        c0 = H[c].x == rm.x ? c : c0;
                        c1 = H[c].y == rm.y ? c : c1;
                    }
                    if (r0) {
                        brow[0] = r;
                        bcol[0] = c0;
                    }
                    if (r1) {
                        brow[1] = r;
                        bcol[1] = c1;
                    }
                }
            }
            best = vmax(best, rm);
            // next row's query code and bias
            const int k = r + 1;
            rb = (q_cur & 0xFFu) * kPairRowBytes;
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
            diagIn = l == 0 ? 0u : bits(inH);
        }
    }

    // Group reduction: global best; with ENDS the earliest row reaching it and
    // the leftmost column of that row.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int mine = h == 0 ? (int)best.x : (int)best.y;
        int top = mine;
#pragma unroll
        for (int d = 1; d < L; d <<= 1) {
            const int o = __shfl_xor(top, d, L);
            top = o > top ? o : top;
        }
        uint32_t key = 0xFFFFFFFFu;
        if constexpr (ENDS) {
            if (top > 0 && top < kMax16 && mine == top && bcol[h] < C)
                key = (uint32_t)brow[h] * (uint32_t)(L * C) + (uint32_t)(l * C + bcol[h]);
#pragma unroll
            for (int d = 1; d < L; d <<= 1) {
                const uint32_t o = (uint32_t)__shfl_xor((int)key, d, L);
                key = o < key ? o : key;
            }
        }
        if (l == 0 && sid[h] != kNone) {
            const uint64_t o = (uint64_t)qid * a.nsub + sid[h];
            if (top >= kMax16) {
                const uint32_t f = atomicAdd(a.nflagged, 1u);
                a.flagged[2 * f] = qid;
                a.flagged[2 * f + 1] = sid[h];
            } else {
                a.score[o] = top;
                if constexpr (ENDS) {
                    uint32_t qe = 0, se = 0;
                    if (top > 0) {
                        qe = key / (uint32_t)(L * C) + 1;
                        se = key % (uint32_t)(L * C) + 1;
                    }
                    a.q_end[o] = qe;
                    a.s_end[o] = se;
                }
            }
        }
    }
}

// Persistent kernel over the shapes of one register class (kFullThreads ==
// NT): each wave claims chunks -- 64 / L subject pairs of one shape against
// one query -- from a single counter until the list is exhausted. The chunk
// list is a sequence of segments, one per shape, each query-rank major
// (queries longest first); a segment's chunk ch of query rank qr is chunk
// start + qr * nch + ch. Every lane of a chunk runs the same query, so the
// same number of steps.
template <bool BIAS, bool ENDS, int NT>
__global__ __launch_bounds__(NT) void full_multi_kernel(FullKernelArgs a) {
    __shared__ uint32_t TP[kPairWords];
    __shared__ FullSeg S[kMaxFullSegs];
    for (int t = threadIdx.x; t < kPairWords; t += NT) TP[t] = a.pair_table[t];
    for (int t = threadIdx.x; t < (int)a.nsegs; t += NT) S[t] = a.segs[t];
    __syncthreads();
    const lds_bytes Tpb = (lds_bytes)TP;
    const int lane = threadIdx.x & 63;
    for (;;) {
        uint32_t ci = 0;
        if (lane == 0) ci = atomicAdd(a.next_chunk, 1u);
        ci = __builtin_amdgcn_readfirstlane(ci);
        if (ci >= a.nchunks) break;
        int k = 0;
        while (k + 1 < (int)a.nsegs && S[k + 1].start <= ci) ++k;
        const uint32_t shape = S[k].shape, nch = S[k].nch, pair_off = S[k].pair_off, npairs = S[k].npairs;
        const uint32_t rel = ci - S[k].start;
        const uint32_t qr = rel / nch, ch = rel - qr * nch;
        const uint32_t qid = a.qorder[qr];
        int id = 0;
#define SABER_FULL_CASE(L_, C_)                                                                 \
        if constexpr (kFullThreads<C_, ENDS> == NT) {                                           \
            if (shape == (uint32_t)id) {                                                        \
                const uint32_t gi = ch * (64 / (L_)) + (uint32_t)lane / (L_);                   \
                if (gi < npairs) full_group<L_, C_, BIAS, ENDS>(a, Tpb, qid, pair_off + gi, lane % (L_)); \
                continue;                                                                       \
            }                                                                                   \
        }                                                                                       \
        ++id;
        if constexpr (ENDS) {
            SABER_FULL_END_SHAPES(SABER_FULL_CASE)
        } else {
            SABER_FULL_SCORE_SHAPES(SABER_FULL_CASE)
        }
#undef SABER_FULL_CASE
    }
}

// Launch and residency of one class kernel (instantiated per flavour in
// full_multi_*.hip).
template <bool BIAS, bool ENDS, int NT>
hipError_t launch_full_class(const FullKernelArgs& a, uint32_t grid, hipStream_t s) {
    SABER_LAUNCH((full_multi_kernel<BIAS, ENDS, NT>), dim3(grid), dim3(NT), 0, s, a);
    return hipGetLastError();
}

// Workgroups of one class kernel resident per CU (asked once; 1 when the
// runtime cannot say, with its error state cleared).
template <bool BIAS, bool ENDS, int NT>
int full_class_blocks_per_cu() {
    static std::atomic<int> cached{0};
    int nb = cached.load(std::memory_order_relaxed);
    if (nb > 0) return nb;
    nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, full_multi_kernel<BIAS, ENDS, NT>, NT, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    nb = nb > 0 ? nb : 1;
    cached.store(nb, std::memory_order_relaxed);
    return nb;
}

}  // namespace saber
