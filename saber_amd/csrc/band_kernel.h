// Packed-int16 banded Smith-Waterman score kernel for gfx950.
//
// Semantics: exactly banded_score (src/algorithm/banded.rs:134-224, 408-449):
// the best score of a local alignment constrained to the diagonal band
// lo <= j - i <= lo + w - 1 and the first row-major cell holding it
// (exclusive ends). The band-cell count of each job depends only on its
// geometry and is filled in by the host planner.
//
// Layout of the work:
//  * Two independent jobs share each lane: every DP value is a pair of int16
//    (job A in the low half, job B in the high half) and every DP operation is
//    one packed VALU op (v_pk_{add,max,min}_i16, v_pk_sub_u16 with clamp).
//  * The DP runs row by row (query position i outer). The band cells of a row
//    live in registers in band coordinates (cell k <-> subject column
//    j = i + lo + k), so the diagonal predecessor is the same register and
//    the vertical predecessor the next one: no shuffles inside a lane.
//  * L lanes cooperate on one job pair, each owning C consecutive band cells.
//    Lane l works on row t - l at step t (a one-row lag per lane), so the
//    only cross-lane values are the left neighbour's E from the previous step
//    and the right neighbour's downward F from the current step, each moved
//    with one DPP row shift.
//  * Gap states are kept clamped at zero: E' = max(E, 0), F' = max(F, 0).
//    With H >= 0 this leaves H = max(0, D, E, F) = max(D, E', F') unchanged and
//    turns each "x - gap" into one unsigned saturating subtract. H - open is
//    formed once per cell and feeds both the next cell's E and the next row's
//    F, which each cell stores as the value leaving it downwards.
//  * Steps run in blocks of four. Each lane keeps its subject columns as a
//    window of bytes that covers a whole block (C + 3 columns), so the byte of
//    cell c at step u of a block is at the fixed position c + u; the window
//    moves by one dword per block. Query codes, bias and new subject columns
//    arrive as dword streams loaded two blocks ahead.
//  * Substitution scores come from an int8 table in LDS (24 query rows x 25
//    columns: the 24 subject codes and a sentinel). Residue buffers are
//    canonical codes (anything >= 24 stored as X) and subject columns outside
//    the window read as the sentinel, which scores -128: a cell left of the
//    rectangle then has a negative diagonal term and zero gap states, so it is
//    exactly 0 (the border value of the scalar kernel), and a cell right of
//    it only receives values from cells before it in row-major order, each
//    lowered (diagonal: by at least 1 even with a +127 bias), so it never
//    exceeds or comes first among the real cells.
//  * Cells beyond the job's clipped width (k >= w) are held at 0: their H and
//    hence their downward F is 0, which is what "outside the band" means for
//    clamped gap states. Only the last K cells of a band can be beyond.
//  * Exactness: every H/E/F of a local alignment is bounded by the optimum; the
//    only upward saturation is H_diag + S + bias > 32767, which forces that
//    cell's H to 32767. A job whose best reaches 32767 is therefore flagged and
//    rescored by the exact int32 kernel; every other job is exact.
//  * End cell: each lane keeps the best of its own cells per job, the row where
//    that best last rose, and a register snapshot of the row's cells taken
//    with one bit-select per cell (v_bfi_b32 on the halves that rose). At the
//    end the group takes the global best, the earliest row that reached it
//    and the leftmost cell of that row.
#pragma once
#include "common.h"

namespace saber {

typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned short v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ v2s vadd(v2s a, v2s b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ v2s vsub(v2s a, v2s b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ v2s vmax(v2s a, v2s b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ v2s vmin(v2s a, v2s b) { return __builtin_elementwise_min(a, b); }
// max(a - b, 0) for a, b in [0, 32767].
__device__ __forceinline__ v2s vsubz(v2s a, v2s b) {
    return __builtin_bit_cast(v2s, __builtin_elementwise_sub_sat(__builtin_bit_cast(v2u, a),
                                                                 __builtin_bit_cast(v2u, b)));
}
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

// A byte stream of one sequence: dword loads at base + o for 32-bit offsets o
// relative to the sequence, clamped into the guard bytes every residue/bias
// buffer carries (16 before the data, 64 after), so each load is valid; bytes
// that do not belong to the sequence are replaced or never used by the caller.
struct ByteStream {
    const uint8_t* base;
    int32_t lo, hi;
    __device__ __forceinline__ void init(const uint8_t* buf, int64_t off, uint64_t len) {
        base = buf + off;
        const int64_t l = -16 - off, h = (int64_t)len + 60 - off;
        lo = (int32_t)(l < INT32_MIN ? INT32_MIN : l);
        hi = (int32_t)(h > INT32_MAX ? INT32_MAX : h);
    }
    __device__ __forceinline__ uint32_t load(int32_t o) const {
        o = o < lo ? lo : (o > hi ? hi : o);
        return gload_u32(base + o);
    }
};

// Subject bytes of window columns j0..j0+3 with columns outside [0, n)
// replaced by the sentinel code. neg8 = -8 * j0 and hi8 = 8 * (n - j0).
__device__ __forceinline__ uint32_t window_dword(uint32_t raw, int32_t neg8, int32_t hi8) {
    const int32_t lo_bits = neg8 < 0 ? 0 : (neg8 > 32 ? 32 : neg8);
    const int32_t hi_bits = hi8 < 0 ? 0 : (hi8 > 32 ? 32 : hi8);
    const uint64_t ones = 0xFFFFFFFFull;
    const uint32_t keep = (uint32_t)(ones << lo_bits) & (uint32_t)((ones << hi_bits) >> 32);
    return (raw & keep) | (kSentinel4 & ~keep);
}

// Pair-table entry offset (bytes) of subject codes a (job A) and b (job B):
// entry a * 25 + b of a query row.
__device__ __forceinline__ uint32_t pair_entries(uint32_t wa, uint32_t wb, int half) {
    // bytes 2*half, 2*half+1 of both windows -> two 16-bit entry offsets
    const uint32_t sel = half ? 0x0C030C02u : 0x0C010C00u;
    const v2u a = __builtin_bit_cast(v2u, __builtin_amdgcn_perm(0u, wa, sel));
    const v2u b = __builtin_bit_cast(v2u, __builtin_amdgcn_perm(0u, wb, sel));
    const v2u e = a * (v2u){100, 100} + (b << (v2u){2, 2});
    return __builtin_bit_cast(uint32_t, e);
}

// The kernel arguments as the workgroup's copy in LDS (a reference to the
// kernel's own parameter would copy it onto every wave's stack).
typedef const BandKernelArgs* band_args;

// The pairs [pair_off, pair_off + npairs) of one chunk, L lanes per pair:
// the group of lanes lane / L scores pair pair_off + lane / L (lane l = lane
// % L of the group). Tb is the int8 score table in LDS; Tp the pair table
// (SHARED only: both jobs of the pair have the same query and first row, so
// a row's query code is common and one LDS read of Tp[q][sA][sB] returns
// both scores already packed). Lane-dependent values are formed here, not in
// the caller, so that the caller keeps no vector registers across the call.
template <int L, int C, int K, bool BIAS, bool SHARED, bool ENDS>
__device__ __attribute__((noinline)) void band_pair(band_args a, const int8_t* Tb, const uint32_t* Tp,
                                                    uint32_t pair_off, uint32_t npairs) {
    const uint32_t lane = __lane_id();
    if (lane / L >= npairs) return;
    const uint32_t g = pair_off + lane / L;
    const int l = (int)(lane % L);
    static_assert(C >= 2 && K >= 1 && K <= C, "variant shape");
    static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "groups must tile a DPP row");
    constexpr int NW = (C + 6) / 4;               // byte window: dwords of 4 columns
    constexpr int NP = ((C + 3 + 3) / 4) * 2;     // pair window: dwords of 2 entries
    constexpr int SPAN = SHARED ? 2 * NP : 4 * NW;  // window columns
    constexpr int NH = SHARED ? 1 : 2;            // query / bias streams
    constexpr uint32_t W = (uint32_t)(L * C);
    const char* Tpb = reinterpret_cast<const char*>(Tp);

    uint32_t jid[2];
    int64_t qoff[2], soff[2];
    uint32_t n[2];
    int lo[2], w[2], ifirst[2], rows[2];
    bool swapped[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        jid[h] = a->pairs[2 * g + h];
        if (jid[h] != kNone) {
            const BandJobDev J = a->jobs[jid[h]];
            swapped[h] = J.swapped;
            qoff[h] = (int64_t)J.q_off;
            soff[h] = (int64_t)J.s_off;
            n[h] = J.n;
            lo[h] = J.lo;
            w[h] = J.w;
            ifirst[h] = J.i_first;
            rows[h] = J.rows;
        } else {
            swapped[h] = false;
            qoff[h] = 0;
            soff[h] = 0;
            n[h] = 0;
            lo[h] = 0;
            w[h] = 1;
            ifirst[h] = 0;
            rows[h] = 0;
        }
    }
    const int R = rows[0] > rows[1] ? rows[0] : rows[1];

    const v2s GO = {(short)a->go, (short)a->go};
    const v2s GE = {(short)a->ge, (short)a->ge};
    const v2s ZERO = {0, 0};

    // Caps for the last K cells of this lane: cell k keeps its value if k < w,
    // else it is held at 0.
    v2s cap[K];
#pragma unroll
    for (int x = 0; x < K; ++x) {
        const int k = l * C + (C - K) + x;
        cap[x] = v2s{k < w[0] ? kMax16 : (short)0, k < w[1] ? kMax16 : (short)0};
    }

    v2s H[C], F[C];
    uint32_t snap[ENDS ? C : 1];   // ENDS: the lane's cells in the row where its best last rose
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H[c] = ZERO;
        F[c] = ZERO;
    }
#pragma unroll
    for (int c = 0; c < (ENDS ? C : 1); ++c) snap[c] = 0;

    // At step t lane l computes row r = t - l; its cell c is window column
    // jb + t + c and its row is query position ifirst + t - l. Block b (steps
    // 4b .. 4b+3) holds window columns jb + 4b .. jb + 4b + SPAN - 1: as bytes
    // (Wd, per job) or as pair-table entry offsets (Wp, both jobs).
    ByteStream ss[2], qs[NH], bs[NH];
    int32_t jb[2], qb[NH], neg8[2], hi8[2];
    uint32_t Wd[SHARED ? 1 : 2][SHARED ? 1 : NW];
    uint32_t Wp[SHARED ? NP : 1];
    uint32_t s_n1[2], s_n2[2], q_cur[NH], q_n1[NH], q_n2[NH];
    uint32_t b_cur[NH], b_n1[NH], b_n2[NH];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        jb[h] = ifirst[h] + lo[h] + l * C - l;
        if (swapped[h])
            ss[h].init(a->qres, soff[h], a->qres_len);
        else
            ss[h].init(a->sres, soff[h], a->sres_len);
        if constexpr (!SHARED) {
#pragma unroll
            for (int x = 0; x < NW; ++x) {
                const int32_t j0 = jb[h] + 4 * x;
                Wd[h][x] = window_dword(ss[h].load(j0), -8 * j0, 8 * ((int32_t)n[h] - j0));
            }
        }
        // Column of the dword appended after block 0, as -8j and 8(n - j).
        neg8[h] = -8 * (jb[h] + SPAN);
        hi8[h] = 8 * ((int32_t)n[h] - jb[h] - SPAN);
        s_n1[h] = ss[h].load(jb[h] + SPAN);
        s_n2[h] = ss[h].load(jb[h] + SPAN + 4);
    }
    if constexpr (SHARED) {
#pragma unroll
        for (int x = 0; x < NP / 2; ++x) {
            uint32_t wv[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int32_t j0 = jb[h] + 4 * x;
                wv[h] = window_dword(ss[h].load(j0), -8 * j0, 8 * ((int32_t)n[h] - j0));
            }
            Wp[2 * x] = pair_entries(wv[0], wv[1], 0);
            Wp[2 * x + 1] = pair_entries(wv[0], wv[1], 1);
        }
    }
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        qb[h] = ifirst[h] - l;
        if (swapped[h])
            qs[h].init(a->sres, qoff[h], a->sres_len);
        else
            qs[h].init(a->qres, qoff[h], a->qres_len);
        q_cur[h] = qs[h].load(qb[h]);
        q_n1[h] = qs[h].load(qb[h] + 4);
        q_n2[h] = qs[h].load(qb[h] + 8);
        b_cur[h] = b_n1[h] = b_n2[h] = 0;
        if constexpr (BIAS) {
            bs[h].init(reinterpret_cast<const uint8_t*>(a->qbias), qoff[h], a->qres_len);
            b_cur[h] = bs[h].load(qb[h]);
            b_n1[h] = bs[h].load(qb[h] + 4);
            b_n2[h] = bs[h].load(qb[h] + 8);
        }
    }

    // Best of this lane's cells (both jobs) and the row where it last rose.
    v2s cbest = ZERO;
    uint32_t brow = 0;
    uint32_t outE = 0;

    const int nblocks = (R + L - 1 + 3) / 4;
    for (int blk = 0; blk < nblocks; ++blk) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * blk + u - l;
            const bool act = (unsigned)r < (unsigned)R;
            uint32_t rb0, rb1 = 0;
            v2s bpair = ZERO;
            if constexpr (SHARED) {
                rb0 = ((q_cur[0] >> (8 * u)) & 0xFFu) * kPairRowBytes;
                if constexpr (BIAS) {
                    const short bv = (short)(int8_t)(b_cur[0] >> (8 * u));
                    bpair = v2s{bv, bv};
                }
            } else {
                rb0 = ((q_cur[0] >> (8 * u)) & 0xFFu) * kTable8Stride;
                rb1 = ((q_cur[NH - 1] >> (8 * u)) & 0xFFu) * kTable8Stride;
                if constexpr (BIAS)
                    bpair = v2s{(short)(int8_t)(b_cur[0] >> (8 * u)), (short)(int8_t)(b_cur[NH - 1] >> (8 * u))};
            }
            v2s inE = ZERO;
            if constexpr (L > 1) {
                inE = pair(from_prev_lane(outE));
                if (l == 0) inE = ZERO;
            }
            v2s E = ZERO, rm = ZERO;
            // One cell: diagonal H[c], E from the left, F arriving from above.
            auto cell = [&](int c, v2s fin) {
                const int p = c + u;
                v2s P;
                if constexpr (SHARED) {
                    const uint32_t e = (Wp[p >> 1] >> (16 * (p & 1))) & 0xFFFFu;
                    uint32_t pv;
                    __builtin_memcpy(&pv, Tpb + rb0 + e, 4);
                    P = pair(pv);
                } else {
                    const uint32_t ca = (Wd[0][p >> 2] >> (8 * (p & 3))) & 0xFFu;
                    const uint32_t cb = (Wd[NH - 1][p >> 2] >> (8 * (p & 3))) & 0xFFu;
                    P = v2s{(short)Tb[rb0 + ca], (short)Tb[rb1 + cb]};
                }
                if constexpr (BIAS) P = vadd(P, bpair);
                const v2s D = vadd(H[c], P);
                v2s hv = vmax(vmax(D, fin), E);
                if (c >= C - K) hv = vmin(hv, cap[c - (C - K)]);
                const v2s hg = vsubz(hv, GO);
                E = vmax(hg, vsubz(E, GE));
                F[c] = vmax(hg, vsubz(fin, GE));
                H[c] = hv;
                rm = vmax(rm, hv);
            };

            if (act) {
                // Cell 0: its left neighbour is the previous lane's last cell.
                E = inE;
                cell(0, F[1]);
            }
            v2s upF = ZERO;
            if constexpr (L > 1) {
                upF = pair(from_next_lane(bits(F[0])));
                if (l == L - 1) upF = ZERO;
            }
            if (act) {
#pragma unroll
                for (int c = 1; c < C - 1; ++c) cell(c, F[c + 1]);
                cell(C - 1, upF);
                outE = bits(E);

                // Rows past a job's own rows do not count for it.
                const v2s live = v2s{r < rows[0] ? kMax16 : (short)0, r < rows[1] ? kMax16 : (short)0};
                rm = vmin(rm, live);
                if constexpr (ENDS) {
                    // Halves whose best rose: all-ones (sign of cbest - rm).
                    uint32_t rose = bits(vsub(cbest, rm) >> 15);
                    // Opaque to the optimiser, so the selects below stay one
                    // v_bfi_b32 per cell instead of per-half selects.
                    asm volatile("" : "+v"(rose));
                    if (rose != 0) {
                        cbest = vmax(cbest, rm);
                        brow = (rose & ((uint32_t)r * 0x10001u)) | (~rose & brow);
#pragma unroll
                        for (int c = 0; c < C; ++c) snap[c] = (rose & bits(H[c])) | (~rose & snap[c]);
                    }
                } else {
                    cbest = vmax(cbest, rm);
                }
            }
        }
        // Next block: the window moves four columns, the streams one dword.
        const int32_t nb = 4 * blk;
        uint32_t wnew[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            wnew[h] = window_dword(s_n1[h], neg8[h], hi8[h]);
            neg8[h] -= 32;
            hi8[h] -= 32;
            s_n1[h] = s_n2[h];
            s_n2[h] = ss[h].load(jb[h] + SPAN + nb + 8);
            if constexpr (!SHARED) {
#pragma unroll
                for (int x = 0; x < NW - 1; ++x) Wd[h][x] = Wd[h][x + 1];
                Wd[h][NW - 1] = wnew[h];
            }
        }
        if constexpr (SHARED) {
#pragma unroll
            for (int x = 0; x < NP - 2; ++x) Wp[x] = Wp[x + 2];
            Wp[NP - 2] = pair_entries(wnew[0], wnew[1], 0);
            Wp[NP - 1] = pair_entries(wnew[0], wnew[1], 1);
        }
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            q_cur[h] = q_n1[h];
            q_n1[h] = q_n2[h];
            q_n2[h] = qs[h].load(qb[h] + nb + 12);
            if constexpr (BIAS) {
                b_cur[h] = b_n1[h];
                b_n1[h] = b_n2[h];
                b_n2[h] = bs[h].load(qb[h] + nb + 12);
            }
        }
    }

    // Group reduction: global best, earliest row reaching it, leftmost cell.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int mine = h == 0 ? (int)cbest.x : (int)cbest.y;
        int best = mine;
#pragma unroll
        for (int d = 1; d < L; d <<= 1) {
            const int o = __shfl_xor(best, d, L);
            best = o > best ? o : best;
        }
        uint32_t key = 0xFFFFFFFFu;
        if (ENDS && best > 0 && best < kMax16 && mine == best) {
            const uint32_t row = h == 0 ? (brow & 0xFFFFu) : (brow >> 16);
            int first = C;
#pragma unroll
            for (int c = C - 1; c >= 0; --c) {
                const uint32_t v = snap[c];
                const int hv = h == 0 ? (int)(short)(v & 0xFFFFu) : (int)(short)(v >> 16);
                first = hv == best ? c : first;
            }
            if (first < C) key = row * W + (uint32_t)(l * C + first);
        }
        if constexpr (ENDS) {
#pragma unroll
            for (int d = 1; d < L; d <<= 1) {
                const uint32_t o = (uint32_t)__shfl_xor((int)key, d, L);
                key = o < key ? o : key;
            }
        }
        if (l == 0 && jid[h] != kNone) {
            const uint32_t id = jid[h];
            if (best >= kMax16) {
                const uint32_t slot = atomicAdd(a->nflagged, 1u);
                a->flagged[slot] = id;
            } else {
                a->out.score[id] = best;
                if constexpr (ENDS) {
                    uint32_t qe = 0, se = 0;
                    if (best > 0) {
                        const uint32_t row = key / W;
                        const uint32_t k = key % W;
                        const int i = ifirst[h] + (int)row;
                        qe = (uint32_t)i + 1;
                        se = (uint32_t)(i + lo[h] + (int)k + 1);
                    }
                    a->out.q_end[id] = qe;
                    a->out.s_end[id] = se;
                }
            }
        }
    }
}

// Every compiled shape (L, C, K), in variant-table order. There is no
// (1, 13, 13): its body with bias, per-job tables and scores only returned
// wrong scores on the device under load (a different value on each launch,
// for most of a batch's jobs of that body; the other flavours of the same
// shape and every other shape were exact), and the cause was not found;
// widths up to 13 run on (1, 9, 9), (1, 17, 8), (1, 17, 17) and (2, 9, 9).
// tests/test_band_stress_gpu.py runs every body and flavour under load.
#define SABER_BAND_SHAPES(X)                                                            \
    X(1, 9, 9) X(1, 17, 8) X(1, 17, 17)                                                 \
    X(2, 9, 9) X(2, 13, 13) X(2, 17, 8) X(2, 17, 17)                                    \
    X(4, 9, 9) X(4, 13, 13) X(4, 17, 8) X(4, 17, 17)                                    \
    X(8, 9, 9) X(8, 13, 13) X(8, 17, 8) X(8, 17, 17)                                    \
    X(16, 9, 9) X(16, 13, 13) X(16, 17, 16) X(16, 17, 17) X(16, 21, 21)                \
    X(2, 33, 2) X(4, 33, 4)

// All shapes in one persistent launch: one workgroup of 12 waves per CU
// stages both score tables in LDS once, then each wave claims work chunks in
// the planner's order (longest first) with one atomic add per chunk until the
// list is exhausted. A chunk is up to 64 lanes of one shape and one flavour:
// the pair-table flavour for pairs that share their query rows, the per-job
// table otherwise. 12 waves per CU = three per SIMD, so every shape must fit
// 168 registers.
template <bool BIAS, bool ENDS>
__global__ __launch_bounds__(768) void band_multi_kernel(BandKernelArgs a) {
    __shared__ uint32_t T4[kTable8Words];
    __shared__ uint32_t TP[kPairWords];
    __shared__ BandKernelArgs SA;
    for (int t = threadIdx.x; t < kTable8Words; t += blockDim.x) T4[t] = a.table[t];
    for (int t = threadIdx.x; t < kPairWords; t += blockDim.x) TP[t] = a.pair_table[t];
    if (threadIdx.x == 0) SA = a;
    __syncthreads();
    const int8_t* Tb = reinterpret_cast<const int8_t*>(T4);
    const band_args ap = &SA;
    const int lane = threadIdx.x & 63;
    for (;;) {
        uint32_t ci = 0;
        if (lane == 0) ci = atomicAdd(a.next_chunk, 1u);
        ci = __builtin_amdgcn_readfirstlane(ci);
        if (ci >= a.nchunks) break;
        // The chunk record in scalar registers: nothing of it is kept in
        // vector registers across the call.
        const BandBlock blk = a.blocks[ci];
        const uint32_t shape = __builtin_amdgcn_readfirstlane(blk.shape);
        const uint32_t pair_off = __builtin_amdgcn_readfirstlane(blk.pair_off);
        const uint32_t npairs = __builtin_amdgcn_readfirstlane(blk.npairs);
        const bool shared = (__builtin_amdgcn_readfirstlane(blk.flags) & kChunkShared) != 0;
        int id = 0;
#define SABER_BAND_CASE(L_, C_, K_)                                                              \
        if (shape == (uint32_t)id) {                                                             \
            if (shared)                                                                          \
                band_pair<L_, C_, K_, BIAS, true, ENDS>(ap, Tb, TP, pair_off, npairs);           \
            else                                                                                 \
                band_pair<L_, C_, K_, BIAS, false, ENDS>(ap, Tb, TP, pair_off, npairs);          \
            continue;                                                                            \
        }                                                                                        \
        ++id;
        SABER_BAND_SHAPES(SABER_BAND_CASE)
#undef SABER_BAND_CASE
    }
}

}  // namespace saber
