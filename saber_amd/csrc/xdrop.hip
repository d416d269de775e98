// Ungapped X-drop extension of seed hits (xdrop_cat, src/search/engine/scan.rs:859-904;
// identical results to ungapped_xdrop_prof, src/search/engine.rs:2425-2455).
//
// For a seed (qp, sp) the right extension scores the pairs (qp+t, sp+t),
// t = 0..nr-1 (the seed pair included), the left extension (qp-t, sp-t),
// t = 1..nl. Each keeps a running sum, its best and the step of the first
// strict improvement, and stops after the step that falls more than x below
// the best. The result is (best_r + best_l, qp - st_l, qp + end_r) and the
// steps of both sides.
//
// One lane runs one hit: the right side, then the left side, then it writes
// the hit's results and takes the next hit. A wave holds a private pool of
// hits (a chunk of consecutive hit indices claimed with one atomic add) and
// hands them to its lanes in lane order, so lanes never wait for a wave's
// longest extension and the shared counter sees a few claims per wave.
//
// Hit ranges are claimed per group of workgroups (see the hit queue below):
// same-address device-scope atomics are serialised, and one counter claimed
// by every wave at start-up and again at the end cost more than the whole
// extension work. Every workgroup must be resident at once (the grid is the
// occupancy limit), since each wave owns its first chunk.
//
// The kernel is bound by the vector memory pipeline (every load or store
// instruction of a wave with scattered addresses costs tens of cycles in the
// texture address and data units), so an iteration issues as few memory
// instructions as possible: the two residue loads of the next visit (issued
// before the scan, consumed in the next iteration) and at most one 16-byte
// result store. Hit descriptors (built by xdrop_desc_kernel) reach the lanes
// through LDS: a wave loads the descriptors of its next 64 hits with one
// coalesced load ahead of time and stages them when the previous 64 are
// handed out. A lane takes its next hit while the current one runs; an idle
// lane starts it, and its first visit runs in the next iteration.
//
// A lane advances 8 steps per visit. The 8 query codes, subject codes and
// bias bytes of the visit come from three aligned dword loads per stream
// (byte-aligned with v_alignbyte, byte-reversed for the left side). Loads run
// one visit ahead: a visit issues the next visit's loads (when the side has
// more than 8 steps left) before it scans, and a hit's start issues both its
// first right visit and its first left visit. Query
// and subject codes are interleaved into 16-bit table indices (q << 8 | s)
// with one v_perm per two steps; the scores come from a 32 x 256 table in
// LDS (int8 when every score fits, else int16). The 8 lookups are issued
// before the sequential scan, which then costs a few VALU per step.
//
// Bytes read beyond an extension's end (at most 11 before a sequence's start
// or 10 after its end) stay inside the 16 guard bytes before and 64 after
// each residue buffer; their scores are never used.
#include "common.h"

namespace saber {

namespace {

// The three aligned dwords holding the 8 bytes of one visit of one stream:
// bytes p..p+7 (forward) or p-7..p (backward).
struct Raw3 {
    uint32_t d0, d1, d2;
};

__device__ __forceinline__ const uint8_t* window_start(const uint8_t* p, bool fwd) { return fwd ? p : p - 7; }

__device__ __forceinline__ Raw3 fetch3(const uint8_t* p, bool fwd) {
    const uint8_t* w0 = window_start(p, fwd);
    // aligned down by pointer arithmetic (keeps the global address space)
    const uint32_t* w = reinterpret_cast<const uint32_t*>(w0 - (reinterpret_cast<uintptr_t>(w0) & 3));
    return Raw3{w[0], w[1], w[2]};
}

// The visit's 8 bytes of a stream, in step order (p, p+1, .. or p, p-1, ..).
__device__ __forceinline__ void decode3(const Raw3& r, const uint8_t* p, bool fwd, uint32_t& s0, uint32_t& s1) {
    const uint32_t off = (uint32_t)(reinterpret_cast<uintptr_t>(window_start(p, fwd)) & 3);
    const uint32_t lo = __builtin_amdgcn_alignbyte(r.d1, r.d0, off);
    const uint32_t hi = __builtin_amdgcn_alignbyte(r.d2, r.d1, off);
    // perm(a, b, sel): selector bytes 0-3 pick b's bytes, 4-7 pick a's.
    s0 = __builtin_amdgcn_perm(hi, lo, fwd ? 0x03020100u : 0x04050607u);
    s1 = __builtin_amdgcn_perm(hi, lo, fwd ? 0x07060504u : 0x00010203u);
}

// Raw bytes of one visit: query, subject and (with a bias) bias streams.
template <bool BIAS>
struct Visit {
    Raw3 q, s, b;
    __device__ __forceinline__ void fetch(const uint8_t* qp, const uint8_t* sp, const int8_t* bp, bool fwd) {
        q = fetch3(qp, fwd);
        s = fetch3(sp, fwd);
        if (BIAS) b = fetch3(reinterpret_cast<const uint8_t*>(bp), fwd);
    }
};

template <bool WIDE>
struct XTable {
    using T = typename std::conditional<WIDE, short, int8_t>::type;
};

}  // namespace

// One hit's start, built once per run by xdrop_desc_kernel: absolute
// positions of the seed pair in the residue buffers, the seed's query
// position and the steps available on each side.
struct alignas(16) XdropDesc {
    uint64_t q_abs, s_abs;
    uint32_t qp, n_right, n_left, pad;
};
static_assert(sizeof(XdropDesc) == 32, "two dwordx4 loads");

__global__ void xdrop_desc_kernel(XdropArgs a) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= a.nhits) return;
    const uint4 r = reinterpret_cast<const uint4*>(a.hits)[h];
    const uint64_t q0 = a.qoff[r.x], s0 = a.soff[r.y];
    const uint32_t m = (uint32_t)(a.qoff[r.x + 1] - q0), sl = (uint32_t)(a.soff[r.y + 1] - s0);
    XdropDesc d;
    d.q_abs = q0 + r.z;
    d.s_abs = s0 + r.w;
    d.qp = r.z;
    d.n_right = min(m - r.z, sl - r.w);   // >= 1: the seed pair
    d.n_left = min(r.z, r.w);
    d.pad = 0;
    reinterpret_cast<XdropDesc*>(a.desc)[h] = d;
}

template <bool BIAS, bool WIDE>
__global__ __launch_bounds__(256) void xdrop_kernel(XdropArgs a) {
    using TT = typename XTable<WIDE>::T;
    __shared__ TT T[kXdropTable];
    __shared__ uint4 Qs[4][64][2];   // per wave: up to 64 staged hit descriptors
    for (int i = threadIdx.x; i < kXdropTable; i += blockDim.x) T[i] = (TT)a.table[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    uint4 (*Q)[2] = Qs[threadIdx.x >> 6];
    const uint4* desc = reinterpret_cast<const uint4*>(a.desc);

    // ---- wave-uniform hit queue. The hits are split into kXdropGroups
    // contiguous ranges, one per group of workgroups (blockIdx mod groups);
    // a group's waves take chunks of its range: chunk j for the group's wave j
    // (no atomic), then further chunks from the group's own counter, so the
    // claims spread over several addresses and none are made at start-up.
    // pool: unclaimed hits of the wave's chunk; LDS holds descriptors of hits
    // qbase .. qbase+qn-1 (qcur handed out); P0/P1 hold the next pn
    // descriptors (hits pbase ..), loaded in advance.
    const uint32_t ngroups = min(kXdropGroups, gridDim.x);
    const uint32_t grp = blockIdx.x % ngroups;
    const uint32_t g_lo = (uint32_t)((uint64_t)a.nhits * grp / ngroups);
    const uint32_t g_hi = (uint32_t)((uint64_t)a.nhits * (grp + 1) / ngroups);
    const uint32_t g_waves = 4 * ((gridDim.x - grp + ngroups - 1) / ngroups);
    const uint32_t g_chunks = (g_hi - g_lo + a.chunk - 1) / a.chunk;
    uint32_t* g_counter = a.next_chunk + grp * kXdropCounterStride;
    uint32_t next_k = 4 * (blockIdx.x / ngroups) + (threadIdx.x >> 6);   // first chunk
    uint32_t pool = 0, pool_end = 0, qbase = 0, qn = 0, qcur = 0, pbase = 0, pn = 0;
    bool exhausted = false;
    uint4 P0{}, P1{};
    auto prefetch = [&]() {
        if (pool == pool_end && !exhausted) {
            if (next_k == UINT32_MAX) {
                // dynamic claim; skip the atomic once the counter is seen past the end
                uint32_t c = 0;
                if (lane == 0) {
                    c = __hip_atomic_load(g_counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (g_waves + c < g_chunks) c = atomicAdd(g_counter, 1u);
                }
                next_k = g_waves + __builtin_amdgcn_readfirstlane(c);
            }
            if (next_k >= g_chunks) {
                exhausted = true;
            } else {
                pool = g_lo + next_k * a.chunk;
                pool_end = min(pool + a.chunk, g_hi);
            }
            next_k = UINT32_MAX;
        }
        pbase = pool;
        pn = min(64u, pool_end - pool);
        pool += pn;
        if (pn) {
            const uint32_t i = pbase + min(lane, pn - 1);
            P0 = desc[2 * i];
            P1 = desc[2 * i + 1];
        }
    };
    prefetch();

    // the lane's next hit, taken from the queue while the current one runs
    bool ready = false;
    uint32_t nh = 0, nqp = 0, nnr = 0, nnl = 0;
    uint64_t nq = 0, ns = 0;

    // the current hit
    bool busy = false, fwd = true, have_v = false, to_left = false, store = false;
    uint32_t h = 0, qp = 0, n = 0, n_left = 0, t = 0, bt = 0, bt_r = 0, t_r = 0;
    int32_t cur = 0, best = 0, best_r = 0;
    uint64_t q_abs = 0, s_abs = 0;
    const uint8_t *qc = nullptr, *sc = nullptr;
    const int8_t* bc = nullptr;
    // bytes of the lane's next visit (loaded by every lane in every iteration,
    // at a dummy address when the lane needs nothing, so that the load always
    // lands in the same registers and is waited for in the next iteration)
    Visit<BIAS> V{};

    for (;;) {
        // ---- this visit's bytes, fetched in the previous iteration
        const bool visit = busy && have_v;
        uint32_t q0w, q1w, s0w, s1w, b0w = 0, b1w = 0;
        decode3(V.q, qc, fwd, q0w, q1w);
        decode3(V.s, sc, fwd, s0w, s1w);
        if (BIAS) decode3(V.b, reinterpret_cast<const uint8_t*>(bc), fwd, b0w, b1w);
        // results of the hit the lane finished in the previous iteration
        // (stored after this iteration's loads are issued)
        const uint4 o_res = make_uint4((uint32_t)(best_r + best), qp - bt, qp + bt_r, t_r + t);
        const uint32_t o_h = h;
        const bool o_store = store;
        store = false;
        // ---- idle lanes start their next hit; a lane done with its right
        // side moves to the left side
        bool start = false;
        if (!busy && ready) {
            h = nh;
            qp = nqp;
            q_abs = nq;
            s_abs = ns;
            n = nnr;
            n_left = nnl;
            fwd = true;
            t = 0;
            cur = 0;
            best = 0;
            bt = 0;
            qc = a.qres + q_abs;
            sc = a.sres + s_abs;
            if (BIAS) bc = a.qbias + q_abs;
            busy = true;
            start = true;
            ready = false;
        }
        if (to_left) {
            fwd = false;
            qc = a.qres + q_abs - 1;
            sc = a.sres + s_abs - 1;
            if (BIAS) bc = a.qbias + q_abs - 1;
            to_left = false;
            start = true;
        }
        // ---- lanes without a next hit take one from the queue
        const uint64_t want = __ballot(!ready);
        if (want) {
            if (qcur == qn && pn) {
                // staged descriptors used up: stage the prefetched ones
                Q[lane][0] = P0;
                Q[lane][1] = P1;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                qbase = pbase;
                qn = pn;
                qcur = 0;
                prefetch();
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
            if (!ready && qcur + rank < qn) {
                const uint4 d0 = Q[qcur + rank][0], d1 = Q[qcur + rank][1];
                nh = qbase + qcur + rank;
                nq = ((uint64_t)d0.y << 32) | d0.x;
                ns = ((uint64_t)d0.w << 32) | d0.z;
                nqp = d1.x;
                nnr = d1.y;
                nnl = d1.z;
                ready = true;
            }
            qcur = min(qcur + (uint32_t)__popcll(want), qn);
        }
        // ---- the next visit's bytes (dummy address for lanes without one)
        const int d = fwd ? 8 : -8;
        const uint32_t rem = n - t;   // >= 1 for a visiting lane
        const bool ahead = visit && rem > 8;   // the side goes on past this visit
        {
            const uint8_t* vq = a.qres;
            const uint8_t* vs = a.qres;
            const int8_t* vb = a.qbias;
            bool vf = true;
            if (ahead) {
                vq = qc + d;
                vs = sc + d;
                if (BIAS) vb = bc + d;
                vf = fwd;
            } else if (start) {
                vq = qc;
                vs = sc;
                if (BIAS) vb = bc;
                vf = fwd;
            }
            V.fetch(vq, vs, vb, vf);
        }
        if (o_store) reinterpret_cast<uint4*>(a.res)[o_h] = o_res;
        have_v = ahead || start;
        if (qcur == qn && pn == 0 && !__ballot(busy || ready)) break;   // nothing left
        if (!visit) continue;

        // ---- scan the visit's 8 steps
        // (q << 8 | s) for steps 0,1 / 2,3 / 4,5 / 6,7, q kept below 32
        const uint32_t p0 = __builtin_amdgcn_perm(q0w, s0w, 0x05010400u) & 0x1FFF1FFFu;
        const uint32_t p1 = __builtin_amdgcn_perm(q0w, s0w, 0x07030602u) & 0x1FFF1FFFu;
        const uint32_t p2 = __builtin_amdgcn_perm(q1w, s1w, 0x05010400u) & 0x1FFF1FFFu;
        const uint32_t p3 = __builtin_amdgcn_perm(q1w, s1w, 0x07030602u) & 0x1FFF1FFFu;
        int32_t v[8];
        v[0] = T[p0 & 0xFFFF];
        v[1] = T[p0 >> 16];
        v[2] = T[p1 & 0xFFFF];
        v[3] = T[p1 >> 16];
        v[4] = T[p2 & 0xFFFF];
        v[5] = T[p2 >> 16];
        v[6] = T[p3 & 0xFFFF];
        v[7] = T[p3 >> 16];
        if (BIAS) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[u] += (int32_t)(int8_t)(((u < 4 ? b0w : b1w) >> (8 * (u & 3))) & 0xFF);
        }
        // all 8 lookups complete before the scan (the compiler would otherwise
        // sink each one behind the previous step's exit test)
#pragma unroll
        for (int u = 0; u < 8; ++u) asm volatile("" : "+v"(v[u]));
        uint32_t took = 8;
        bool done = false;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cur += v[u];
            if (cur > best) {
                best = cur;
                bt = t + u + 1;
            }
            if (cur < best - a.x || (uint32_t)(u + 1) >= rem) {
                took = u + 1;
                done = true;
                break;
            }
        }
        t += took;
        if (!done) {
            qc += d;
            sc += d;
            if (BIAS) bc += d;
            continue;
        }
        have_v = false;
        if (fwd) {
            best_r = best;
            bt_r = bt;
            t_r = t;
            n = n_left;
            t = 0;
            cur = 0;
            best = 0;
            bt = 0;
            if (n > 0) {
                to_left = true;
                continue;
            }
        }
        busy = false;
        store = true;
    }
    if (store)
        reinterpret_cast<uint4*>(a.res)[h] = make_uint4((uint32_t)(best_r + best), qp - bt, qp + bt_r, t_r + t);
}

int xdrop_blocks_per_cu(bool bias, bool wide) {
    int n = 0;
    hipError_t e;
    if (bias)
        e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xdrop_kernel<true, true>, 256, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xdrop_kernel<true, false>, 256, 0);
    else
        e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xdrop_kernel<false, true>, 256, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xdrop_kernel<false, false>, 256, 0);
    return e == hipSuccess && n > 0 ? n : 1;
}

hipError_t launch_xdrop(const XdropArgs& a, bool bias, bool wide, int grid, hipStream_t s) {
    if (a.nhits == 0) return hipSuccess;
    SABER_LAUNCH(xdrop_desc_kernel, dim3((a.nhits + 255) / 256), dim3(256), 0, s, a);
    return launch_xdrop_described(a, bias, wide, grid, s);
}

hipError_t launch_xdrop_described(const XdropArgs& a, bool bias, bool wide, int grid, hipStream_t s) {
    if (a.nhits == 0) return hipSuccess;
    if (bias) {
        if (wide) SABER_LAUNCH((xdrop_kernel<true, true>), dim3(grid), dim3(256), 0, s, a);
        else SABER_LAUNCH((xdrop_kernel<true, false>), dim3(grid), dim3(256), 0, s, a);
    } else {
        if (wide) SABER_LAUNCH((xdrop_kernel<false, true>), dim3(grid), dim3(256), 0, s, a);
        else SABER_LAUNCH((xdrop_kernel<false, false>), dim3(grid), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace saber
