// Banded Smith-Waterman with traceback on the device: banded_align_until
// (src/algorithm/banded.rs:472-521), i.e. the scalar recurrence of
// banded_align_scalar (banded.rs:227-333) with its direction bits
// (org | e_ext << 2 | f_ext << 3: origin 0 start, 1 diagonal, 2 E, 3 F; ties
// diagonal > E > F, extension only when strictly better), the row stop of the
// AVX2 kernels the dispatcher runs (banded_simd.rs:324, 594: after the first
// row holding band cells whose running best reaches stop_at), and walk
// (banded.rs:336-403).
//
// Band widths up to 64 * 17 run L lanes of C cells per job (the
// lane-parallel kernel below, direction rows of L C bytes), wider ones one
// thread per job (rows of w bytes); the direction bytes of a job's band rows
// (rows i_first .. i_first + rows - 1) go to a per-job slice of a direction
// buffer.
// The walk merges equal operations into runs as it goes (in registers),
// writes the runs in reverse order at the job's offset of the output arrays
// and then reverses them in place: no value it stores is read back while
// walking.
#include "common.h"

namespace saber {

constexpr int32_t kNegAlign = INT32_MIN / 4;

__device__ __forceinline__ int32_t amax(int32_t a, int32_t b) { return a > b ? a : b; }

// Jobs wider than the lane-parallel form (below): one thread per job, the two
// band rows updated in place in a per-job slice of a work buffer (a cell
// reads the previous row's H[k], H[k+1], F[k+1] before it writes H[k], F[k];
// cells below the row's first band cell are never valid -- the first cell
// only moves left from row to row -- and cells past its last are reset after
// the row); cells four at a time, the previous row's values and one dword of
// subject codes loaded together (the codes one group ahead; the residue
// buffers carry guard bytes past their end).
constexpr int kWideThreads = 64;

__device__ void write_walk(const AlignArgs& a, uint32_t slot, const BandJobDev& J, const uint8_t* dir, int64_t stride,
                           int32_t best, int64_t bi, int64_t bk, AlignRecord rec) {
    const uint32_t J_id = a.ids[slot];
    const int64_t lo = J.lo;
    uint8_t* oc = a.op_codes + a.op_off[slot];
    uint32_t* ol = a.op_lens + a.op_off[slot];
    // walk from the end cell backwards; runs of equal operations in reverse
    uint32_t nr = 0, run = 0;
    uint8_t cur = 0;
    int64_t nq = 0, ns = 0;   // query / window residues the operations consume
    auto emit = [&](uint8_t op) {
        if (op == cur) {
            ++run;
            return;
        }
        if (run) {
            oc[nr] = cur;
            ol[nr] = run;
            ++nr;
        }
        cur = op;
        run = 1;
    };
    int64_t i = bi, k = bk;
    int state = 0;   // 0 = H, 1 = E (query gap, 'D'), 2 = F (subject gap, 'I')
    for (;;) {
        const uint8_t d = dir[(i - J.i_first) * stride + k];
        if (state == 0) {
            const uint32_t org = d & 3u;
            if (org == 0) break;
            if (org == 1) {
                emit('M');
                ++nq;
                ++ns;
                if (i == 0 || i + lo + k == 0) break;
                i -= 1;
            } else {
                state = org == 2 ? 1 : 2;
            }
        } else if (state == 1) {
            emit('D');
            ++ns;
            k -= 1;
            if (!(d & 4u)) state = 0;
        } else {
            emit('I');
            ++nq;
            i -= 1;
            k += 1;
            if (!(d & 8u)) state = 0;
        }
    }
    if (run) {
        oc[nr] = cur;
        ol[nr] = run;
        ++nr;
    }
    // alignment order
    for (int64_t x = 0, y = (int64_t)nr - 1; x < y; ++x, --y) {
        const uint8_t cx = oc[x], cy = oc[y];
        const uint32_t lx = ol[x], ly = ol[y];
        oc[x] = cy;
        oc[y] = cx;
        ol[x] = ly;
        ol[y] = lx;
    }
    const int64_t qe = bi + 1, se = bi + lo + bk + 1;
    rec.status = 0;
    rec.score = best;
    rec.q_start = (uint32_t)(qe - nq);
    rec.q_end = (uint32_t)qe;
    rec.s_start = (uint32_t)(se - ns);
    rec.s_end = (uint32_t)se;
    rec.nops = nr;
    a.out[J_id] = rec;
}

__global__ __launch_bounds__(kWideThreads) void band_align_kernel(AlignArgs a) {
    __shared__ int32_t M[576];
    for (int x = threadIdx.x; x < 576; x += blockDim.x) M[x] = a.mat[x];
    __syncthreads();
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
    int32_t* H = a.work + a.work_off[t];
    int32_t* F = H + (w + 1);
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
        const int32_t* row = M + qc * 24;
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
    write_walk(a, t, J, dir, w, best, bi, bk, rec);
}

// Lane-parallel form for band widths up to 64 * 17: L lanes per job of C
// cells each (17, or 5 for long jobs), lane l owning band cells
// l*C .. l*C + C - 1 and working on row r = t - l at step t
// (the scoring kernel's layout, in exact int32 with NEG outside the band). A
// cell's diagonal and upper neighbours are the lane's own previous-row values
// except the last cell's upper-right one, which is the right lane's first
// cell of the row above -- computed at the same step, so every lane forms its
// first cell first and the values cross lanes once, by DPP -- and the first
// cell's left neighbour is the left lane's last cell of the same row (from the
// previous step). Each lane keeps its first row-major best (value, row,
// cell) and, for the row stop, a snapshot of it taken at the end of the first
// row in which it reached stop_at: the row stop (first row holding a cell >=
// stop_at) is the least such row over the lanes, and the best over the rows
// up to it is among the snapshots of the lanes that reached it there. After a
// device-scope fence, the group's first lane walks the direction bytes.
template <int L>
__device__ __forceinline__ int32_t lane_from_prev(int32_t v) {
    if constexpr (L <= 16)
        return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);   // row_shr:1
    else
        return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false);   // wave_shr:1
}
template <int L>
__device__ __forceinline__ int32_t lane_from_next(int32_t v) {
    if constexpr (L <= 16)
        return __builtin_amdgcn_update_dpp(0, v, 0x101, 0xF, 0xF, false);   // row_shl:1
    else
        return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false);   // wave_shl:1
}

// One job on the L lanes of a group: launch slot g, lane l (M: the
// substitution matrix in LDS).
template <int L, int C>
__device__ __forceinline__ void align_lanes_job(const AlignArgs& a, const int32_t* M, uint32_t g, int l) {
    static_assert(C == 5 || C == 17, "lane cells");
    constexpr int NW = (C + 3) / 4;   // dwords of a lane's subject codes / direction bytes
    const uint32_t id = a.ids[g];
    const BandJobDev J = a.jobs[id];
    const int32_t w = J.w, lo = J.lo, n = (int32_t)J.n, rows = J.rows, i_first = J.i_first;
    constexpr int32_t W = L * C;   // direction-row stride: a lane's C bytes are contiguous
    const int32_t go = a.go, ge = a.ge;
    const int32_t stop = a.stop_at ? a.stop_at[id] : INT32_MAX;
    const uint8_t* q = a.qres + J.q_off;
    const uint8_t* s = a.sres + J.s_off;
    // The bias load is unconditional (from the query codes when there is no
    // bias, the value then masked), and so is the direction store (steps
    // outside the job's rows write a spare row after them): the number of
    // memory operations per step is fixed, so waiting for the next row's codes
    // does not wait for this row's store.
    const bool has_bias = a.qbias != nullptr;
    const int8_t* bias = has_bias ? a.qbias + J.q_off : reinterpret_cast<const int8_t*>(q);
    uint8_t* dir = a.dir + a.dir_off[g];
    int32_t H[C], F[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H[c] = kNegAlign;
        F[c] = kNegAlign;
    }
    int32_t outE = kNegAlign, outH = kNegAlign;   // this lane's last cell, for the right lane
    int32_t best = 0, bi = 0, bk = 0;             // first row-major best of this lane
    int32_t sbest = 0, si = 0, sk = 0, srow = INT32_MAX;   // snapshot at the first row reaching stop
    // subject bytes of a row: columns i + lo + l*C .. + C - 1, as NW
    // unaligned dwords clamped into the buffer's guard bytes (bytes outside
    // the window belong to cells outside the band rectangle, never used)
    auto load_cols = [&](int32_t j0, uint32_t (&v)[NW]) {
#pragma unroll
        for (int x = 0; x < NW; ++x) {
            int32_t o = j0 + 4 * x;
            o = o < -16 ? -16 : (o > n + 60 ? n + 60 : o);
            __builtin_memcpy(&v[x], s + o, 4);
        }
    };
    uint32_t sv[NW], svn[NW];
    load_cols(i_first + lo + l * C - l, svn);
    // query code and bias of the next row, loaded a step ahead (index
    // clamped into the query)
    const int32_t m = (int32_t)J.m;
    auto qrow = [&](int32_t i, uint32_t& qc, int32_t& b) {
        const int32_t ic = i < 0 ? 0 : (i >= m ? m - 1 : i);
        const uint8_t qv = q[ic];
        const int32_t bv = bias[ic];
        qc = qv < 24 ? qv : (uint32_t)kCodeX;
        b = has_bias ? bv : 0;
    };
    uint32_t qc_next;
    int32_t b_next;
    qrow(i_first - l, qc_next, b_next);
    const int32_t nsteps = rows + L - 1;
    for (int32_t t = 0; t < nsteps; ++t) {
        const int32_t r = t - l;
        const bool act = r >= 0 && r < rows;
        const int32_t i = i_first + r;
#pragma unroll
        for (int x = 0; x < NW; ++x) sv[x] = svn[x];
        load_cols(i + 1 + lo + l * C, svn);   // next row's columns
        // left lane's last cell of this row (previous step)
        int32_t inE = lane_from_prev<L>(outE), inH = lane_from_prev<L>(outH);
        if (l == 0) {
            inE = kNegAlign;
            inH = kNegAlign;
        }
        const uint32_t qc = qc_next;
        const int32_t b = b_next;
        qrow(i + 1, qc_next, b_next);
        int32_t kmin = 0, kmax = -1;
        if (act) {
            kmin = -(i + lo);
            if (kmin < 0) kmin = 0;
            kmax = n - 1 - i - lo;
            if (kmax > w - 1) kmax = w - 1;
        }
        // the row's scores, read from LDS together before the cell chain
        const int32_t* row = M + qc * 24;
        int32_t sc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const uint32_t code = (sv[c >> 2] >> (8 * (c & 3))) & 0xFFu;
            sc[c] = row[code < 24 ? code : (uint32_t)kCodeX] + b;
        }
        uint32_t d4[NW] = {};   // the lane's direction bytes of this row
        int32_t e = inE, hl = inH;
        bool has_cells = false;   // a band cell of this lane in this row
        const bool up_ok = i >= 1;
        // one cell: old H/F of the cell (diagonal) and of the next one (up);
        // branch-free: a cell outside the band rectangle takes NEG on an
        // active row and keeps its values on a lag step
        auto cell = [&](int c, int32_t hup, int32_t fup) {
            const int32_t k = l * C + c;
            const bool valid = act && k >= kmin && k <= kmax;
            const int32_t j = i + lo + k;
            const int32_t eo = hl - go, ee = e - ge;
            const uint32_t e_ext = ee > eo;
            const int32_t ev = k > kmin ? amax(ee, eo) : kNegAlign;
            const int32_t fo = hup - go, fe = fup - ge;
            const uint32_t f_ext = fe > fo;
            const int32_t f = up_ok ? amax(fe, fo) : kNegAlign;
            const int32_t hd = (up_ok && j >= 1) ? H[c] : 0;
            const int32_t diag = hd + sc[c];
            const uint32_t c1 = diag > 0;
            const int32_t h1 = amax(diag, 0);
            const uint32_t c2 = ev > h1;
            const int32_t h2 = amax(h1, ev);
            const uint32_t c3 = f > h2;
            const int32_t h = amax(h2, f);
            const uint32_t org = c3 ? 3u : (c2 ? 2u : c1);
            const int32_t hv = valid ? h : kNegAlign;
            H[c] = act ? hv : H[c];
            F[c] = act ? (valid ? f : kNegAlign) : F[c];
            e = act ? (valid ? ev : kNegAlign) : e;
            hl = act ? hv : hl;
            d4[c >> 2] |= (valid ? (org | (e_ext << 2) | (f_ext << 3)) : 0u) << (8 * (c & 3));
            const bool better = hv > best;   // hv is NEG unless valid
            best = better ? hv : best;
            bi = better ? i : bi;
            bk = better ? k : bk;
            has_cells = has_cells || valid;
        };
        // first cell, then its new values to the left lane (the cell above-
        // right of that lane's last cell)
        cell(0, H[1], F[1]);
        int32_t upH = lane_from_next<L>(H[0]), upF = lane_from_next<L>(F[0]);
        if (l == L - 1) {
            upH = kNegAlign;
            upF = kNegAlign;
        }
#pragma unroll
        for (int c = 1; c < C - 1; ++c) cell(c, H[c + 1], F[c + 1]);
        cell(C - 1, upH, upF);
        // C bytes at their cells' offsets: whole dwords (unaligned), then
        // the last byte
        uint8_t* drow = dir + (int64_t)(act ? r : rows) * W + l * C;
        __builtin_memcpy(drow, d4, 4 * (NW - 1));
        drow[C - 1] = (uint8_t)d4[NW - 1];
        if (act) {
            outE = e;
            outH = hl;
            if (srow == INT32_MAX && best >= stop && has_cells) {
                srow = i;
                sbest = best;
                si = bi;
                sk = bk;
            }
        }
    }
    // the row stop and the group's best
    int32_t stop_row = srow;
#pragma unroll
    for (int d = 1; d < L; d <<= 1) {
        const int32_t o = __shfl_xor(stop_row, d, L);
        stop_row = o < stop_row ? o : stop_row;
    }
    int32_t cv, ci, ck;
    if (stop_row != INT32_MAX) {
        const bool mine = srow == stop_row;
        cv = mine ? sbest : INT32_MIN;
        ci = si;
        ck = sk;
    } else {
        cv = best;
        ci = bi;
        ck = bk;
    }
    int32_t top = cv;
#pragma unroll
    for (int d = 1; d < L; d <<= 1) {
        const int32_t o = __shfl_xor(top, d, L);
        top = o > top ? o : top;
    }
    uint32_t key = (cv == top && top > 0) ? (uint32_t)(ci - i_first) * (uint32_t)W + (uint32_t)ck : 0xFFFFFFFFu;
#pragma unroll
    for (int d = 1; d < L; d <<= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)key, d, L);
        key = o < key ? o : key;
    }
    __threadfence();
    if (l != 0) return;
    // cells of the rows computed by the reference (up to the row stop)
    const int64_t last_row = stop_row != INT32_MAX ? (int64_t)stop_row : i_first + rows - 1;
    uint64_t cells = 0;
    for (int64_t i = i_first; i <= last_row; ++i) {
        int64_t kmin = -(i + lo);
        if (kmin < 0) kmin = 0;
        int64_t kmax = n - 1 - i - lo;
        if (kmax > w - 1) kmax = w - 1;
        cells += (uint64_t)(kmax - kmin + 1);
    }
    AlignRecord rec{};
    rec.status = 1;
    rec.cells = cells;
    if (top <= 0) {
        a.out[id] = rec;
        return;
    }
    const int64_t bi0 = i_first + key / (uint32_t)W, bk0 = key % (uint32_t)W;
    write_walk(a, g, J, dir, W, top, bi0, bk0, rec);
}

// Every lane class in one launch: class x (L = 1 << x lanes per job, 256 / L
// jobs a block) has block_start[x + 1] - block_start[x] blocks, its jobs at
// launch slots slot_start[x] .. slot_start[x + 1] - 1; the block order table
// maps launch blocks to (class, block of the class), longest jobs first.
__global__ __launch_bounds__(256) void band_align_lanes_kernel(AlignArgs a, AlignLaneClasses cls) {
    __shared__ int32_t M[576];
    for (int x = threadIdx.x; x < 576; x += blockDim.x) M[x] = a.mat[x];
    __syncthreads();
    const uint32_t e = cls.order[blockIdx.x];
    const int x = (int)(e >> 24);
    const uint32_t lb = e & 0xFFFFFFu;
    const uint32_t end = cls.slot_start[x + 1];
    switch (x) {
#define SABER_ALIGN_CLASS(X)                                                                         \
    case X: {                                                                                        \
        constexpr int L = align_class_lanes(X), C = align_class_cells(X);                            \
        const uint32_t g = cls.slot_start[X] + (lb * 256 + threadIdx.x) / L;                        \
        if (g < end) align_lanes_job<L, C>(a, M, g, (int)(threadIdx.x % L));                        \
        break;                                                                                       \
    }
        SABER_ALIGN_CLASS(0)
        SABER_ALIGN_CLASS(1)
        SABER_ALIGN_CLASS(2)
        SABER_ALIGN_CLASS(3)
        SABER_ALIGN_CLASS(4)
        SABER_ALIGN_CLASS(5)
        SABER_ALIGN_CLASS(6)
        SABER_ALIGN_CLASS(7)
        SABER_ALIGN_CLASS(8)
        SABER_ALIGN_CLASS(9)
        SABER_ALIGN_CLASS(10)
        SABER_ALIGN_CLASS(11)
#undef SABER_ALIGN_CLASS
    default: break;
    }
}

// The operations of one launch packed densely for the copy to the host:
// slot x's nops entries move from op_off[x] to dst_off[x] (the exclusive
// prefix of nops over the launch's slots that aligned). One wave per slot.
__global__ __launch_bounds__(256) void gather_ops_kernel(uint32_t n, const uint32_t* ids, const AlignRecord* recs,
                                                         const uint64_t* op_off, const uint64_t* dst_off,
                                                         const uint8_t* codes, const uint32_t* lens,
                                                         uint8_t* dcodes, uint32_t* dlens) {
    const uint32_t x = blockIdx.x * 4 + threadIdx.x / 64;
    if (x >= n) return;
    const AlignRecord& r = recs[ids[x]];
    if (r.status != 0) return;
    const uint64_t src = op_off[x], dst = dst_off[x];
    for (uint32_t k = threadIdx.x % 64; k < r.nops; k += 64) {
        dcodes[dst + k] = codes[src + k];
        dlens[dst + k] = lens[src + k];
    }
}

hipError_t launch_gather_ops(uint32_t n, const uint32_t* ids, const AlignRecord* recs, const uint64_t* op_off,
                             const uint64_t* dst_off, const uint8_t* codes, const uint32_t* lens, uint8_t* dcodes,
                             uint32_t* dlens, hipStream_t s) {
    if (n == 0) return hipSuccess;
    SABER_LAUNCH(gather_ops_kernel, dim3((n + 3) / 4), dim3(256), 0, s, n, ids, recs, op_off, dst_off, codes,
                       lens, dcodes, dlens);
    return hipGetLastError();
}

hipError_t launch_band_align(const AlignArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    SABER_LAUNCH(band_align_kernel, dim3((a.n + kWideThreads - 1) / kWideThreads), dim3(kWideThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_band_align_lanes(const AlignArgs& a, const AlignLaneClasses& cls, hipStream_t s) {
    const uint32_t blocks = cls.block_start[kAlignLaneClasses];
    if (blocks == 0) return hipSuccess;
    SABER_LAUNCH(band_align_lanes_kernel, dim3(blocks), dim3(256), 0, s, a, cls);
    return hipGetLastError();
}

}  // namespace saber
