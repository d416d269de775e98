// Learned region prefilter: a 64 -> 256 -> 256 -> 1 MLP (ReLU) with bf16
// weights and activations and fp32 accumulation, trained and run by
// hand-written bf16 MFMA kernels (v_mfma_f32_16x16x32_bf16).
//
// Layers are computed in transposed form, D^T = W^T . X^T, so the weights are
// the MFMA A operand and the activations the B operand. In the 16x16x32
// instruction a lane holds, for operand A, row (lane % 16) and the 8 k values
// 8*(lane/16) .. +7; for B, column (lane % 16) and the same k values; the
// result D holds column (lane % 16), rows 4*(lane/16) .. +3. A layer-1 result
// tile therefore gives each lane one row of the batch and 4 consecutive hidden
// units, and two adjacent tiles give the 8 k values layer 2 needs from that
// lane -- in a permuted k order. Permuting the matching k index of the next
// layer's weights (W2P, W2Q below) absorbs the permutation, so activations
// never leave registers between layers.
//
// k permutation: for a 32-wide k step s and lane group g (= lane / 16), slot
// t' in 0..7 of the operand holds unit 32*s + 4*g + t' (t' < 4) or
// 32*s + 16 + 4*g + (t' - 4) (t' >= 4).
#include <algorithm>

#include "common.h"

namespace saber {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 64, kHid = 256;

__device__ __forceinline__ int perm_unit(int s, int g, int t) {
    return 32 * s + (t < 4 ? 4 * g + t : 16 + 4 * g + (t - 4));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) {
    bf16x8 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

// Weights as the kernels read them (bf16, device):
//   W1T [256][64]   W1T[h][i] = W1[i][h]           (layer-1 A operand)
//   W2P [256][256]  W2P[o][s*32 + 8g + t] = W2[perm_unit(s, g, t)][o]   (layer-2 A)
//   W2Q [256][256]  W2Q[h][s*32 + 8g + t] = W2[h][perm_unit(s, g, t)]   (backward A)
//   b1, b2, w3 fp32 [256], b3 fp32 [1].

// Forward for one 16-row tile held by one wave. Produces the logit of the
// lane's row (valid in every lane; rows past n are computed on zeros) and,
// when `keep`, the post-ReLU H1 and H2 in D layout.
struct Tile {
    f32x4 h1[16];  // layer-1 post-ReLU, tile ht: units 16*ht + 4g + t
    f32x4 h2[16];  // layer-2 post-ReLU, tile ot: units 16*ot + 4g + t
};

template <bool KEEP>
__device__ __forceinline__ float forward_tile(const MlpWeights& w, const __bf16* X, uint64_t row0,
                                              uint64_t n, int lane, Tile* keep) {
    const int g = lane >> 4, r = lane & 15;
    const uint64_t row = row0 + r;
    // X^T fragments: lane holds X[row][32*ks + 8g .. +7]
    bf16x8 xb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        if (row < n)
            xb[ks] = load8(X + row * kIn + 32 * ks + 8 * g);
        else
            xb[ks] = bf16x8{};
    }
    float y = 0.f;
    f32x4 acc2[16];
#pragma unroll
    for (int ot = 0; ot < 16; ++ot) acc2[ot] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        // layer-1 tiles 2s and 2s+1 -> layer-2 operand for k step s
        f32x4 a1[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int ht = 2 * s + u;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mfma(load8(w.W1T + (uint64_t)(16 * ht + r) * kIn + 32 * ks + 8 * g), xb[ks], acc);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float v = acc[t] + w.b1[16 * ht + 4 * g + t];
                acc[t] = v > 0.f ? v : 0.f;
            }
            a1[u] = acc;
            if constexpr (KEEP) keep->h1[ht] = acc;
        }
        bf16x8 hb;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            hb[t] = (__bf16)a1[0][t];
            hb[4 + t] = (__bf16)a1[1][t];
        }
#pragma unroll
        for (int ot = 0; ot < 16; ++ot)
            acc2[ot] = mfma(load8(w.W2P + (uint64_t)(16 * ot + r) * kHid + 32 * s + 8 * g), hb, acc2[ot]);
    }
#pragma unroll
    for (int ot = 0; ot < 16; ++ot) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int u = 16 * ot + 4 * g + t;
            float v = acc2[ot][t] + w.b2[u];
            v = v > 0.f ? v : 0.f;
            acc2[ot][t] = v;
            y += v * w.w3[u];
        }
        if constexpr (KEEP) keep->h2[ot] = acc2[ot];
    }
    // rows are lane % 16: sum the four lane groups
    y += __shfl_xor(y, 16);
    y += __shfl_xor(y, 32);
    return y + *w.b3;
}

// Inference: both weight matrices live in LDS for the whole launch (W1T 32 KB
// + W2P 128 KB, the 160 KB a CU holds), staged once per block, one block per
// CU. Every wave then runs whole row groups on its own -- no barrier after the
// staging: R 16-row tiles per group, layer 1 from W1T fragments, its ReLU'd
// bf16 results kept in registers as layer 2's B operands (the permuted k
// order of W2P), layer 2 in chunks of OC output tiles with the ReLU . w3
// epilogue. Each A fragment read from LDS feeds R MFMAs. Biases initialise
// the accumulators (MFMA C operand).
//
// LDS layout, chunk-column major: 16-byte chunk c (k values 8c .. 8c+7) of
// weight row `row` sits at (c * 256 + row) * 16, W2P after W1T. The 16 lanes
// of a lane group read rows 16*tile .. +15 of one chunk column -- 256
// contiguous bytes, free of bank conflicts -- and every fragment address is
// the lane's base plus a compile-time offset.
constexpr uint32_t kW1Bytes = kHid * kIn * 2, kW2Bytes = kHid * kHid * 2;
constexpr uint32_t kChunkCol = 256 * 16;   // bytes of one chunk column

__device__ __forceinline__ bf16x8 lds8(const char* base, uint32_t off) {
    return *reinterpret_cast<const bf16x8*>(base + off);
}

// Two fp32 values -> ReLU'd bf16 pair in one dword (low half = a). ReLU is
// taken on the bf16 bit patterns as signed 16-bit integers: negative values
// (sign bit set) become +0, the rest are kept -- the same as rounding
// max(x, 0).
__device__ __forceinline__ uint32_t relu_pack(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef short s16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const bf16x2 h = __builtin_convertvector((f32x2{a, b}), bf16x2);
    s16x2 v;
    __builtin_memcpy(&v, &h, 4);
    const s16x2 z = {0, 0};
    v = v > z ? v : z;
    uint32_t out;
    __builtin_memcpy(&out, &v, 4);
    return out;
}

// max(x, 0) as a signed integer max on the bit pattern (exact for every
// non-NaN x; one instruction, no canonicalisation)
__device__ __forceinline__ float relu(float x) {
    return __int_as_float(max(__float_as_int(x), 0));
}

template <int R, int OC, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void mlp_forward_kernel(MlpWeights w, const __bf16* X, uint64_t n,
                                                                 float* logits) {
    __shared__ __attribute__((aligned(16))) char lds[kW1Bytes + kW2Bytes];
    // staging: consecutive threads fill consecutive LDS chunks (one chunk
    // column, consecutive rows)
    for (uint32_t i = threadIdx.x; i < (kW1Bytes + kW2Bytes) / 16; i += blockDim.x) {
        const uint32_t c = i >> 8, row = i & 255;
        const uint4 v = c < 8 ? reinterpret_cast<const uint4*>(w.W1T)[row * 8 + c]
                              : reinterpret_cast<const uint4*>(w.W2P)[row * 32 + (c - 8)];
        *reinterpret_cast<uint4*>(lds + 16 * i) = v;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    const uint64_t groups = (n + 16 * R - 1) / (16 * R);
    const float b3 = *w.b3;
    // Biases held across the wave: lane (g, r) holds units 16r + 4g .. +3 of
    // b1, b2 and w3; the accumulators of a tile take them from lane (g, tile)
    // by ds_bpermute (an LDS-pipe shuffle, no memory access), which keeps the
    // vector-memory counter for the input rows alone.
    const f32x4 vb1 = *reinterpret_cast<const f32x4*>(w.b1 + 16 * r + 4 * g);
    const f32x4 vb2 = *reinterpret_cast<const f32x4*>(w.b2 + 16 * r + 4 * g);
    const f32x4 vw3 = *reinterpret_cast<const f32x4*>(w.w3 + 16 * r + 4 * g);
    // input rows of a group (rows past n read row n - 1; their logits are not
    // stored), loaded one group ahead
    auto load_x = [&](uint64_t grp, bf16x8 (&xb)[R][2]) {
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint64_t row = min(grp * 16 * R + 16 * t + r, n - 1);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) xb[t][ks] = load8(X + row * kIn + 32 * ks + 8 * g);
        }
    };
    uint64_t grp = (uint64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    bf16x8 xn[R][2];
    load_x(grp, xn);
    for (; grp < groups; grp += nwaves) {
        const uint64_t row0 = grp * 16 * R;
        bf16x8 xb[R][2];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            xb[t][0] = xn[t][0];
            xb[t][1] = xn[t][1];
        }
        load_x(grp + nwaves, xn);
        // Loop-invariant addresses are made opaque inside the loop, so that
        // each fragment or shuffle address is a base plus an offset that fits
        // the instruction's immediate field (rather than one register per
        // address, hoisted out of the loop): the lane's LDS base (chunk
        // column g, row r) and the shuffle base (lane 16g).
        uint32_t lb = (uint32_t)g * kChunkCol + (uint32_t)r * 16u, sh = 64u * g;
        asm volatile("" : "+v"(lb), "+v"(sh));
        auto bcast = [&](const f32x4& v, int tile) {
            f32x4 o;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                o[e] = __int_as_float(__builtin_amdgcn_ds_bpermute((int)(sh + 4 * tile), __float_as_int(v[e])));
            return o;
        };
        const char* L1 = lds + lb;
        // layer 1: tiles 2s, 2s+1 form layer 2's operand for k step s; the
        // next tile's fragments are read while this tile's MFMAs run
        uint32_t hw[R][8][4];   // bf16 pairs: k step s, dwords 0-1 tile 2s, 2-3 tile 2s+1
        bf16x8 a0 = lds8(L1, 0), a1 = lds8(L1, 4 * kChunkCol);
#pragma unroll
        for (int ht = 0; ht < 16; ++ht) {
            bf16x8 n0, n1;
            if (ht < 15) {
                n0 = lds8(L1, 256 * (ht + 1));
                n1 = lds8(L1, 4 * kChunkCol + 256 * (ht + 1));
            }
            const f32x4 bias = bcast(vb1, ht);
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const f32x4 acc = mfma(a1, xb[t][1], mfma(a0, xb[t][0], bias));
                hw[t][ht >> 1][2 * (ht & 1)] = relu_pack(acc[0], acc[1]);
                hw[t][ht >> 1][2 * (ht & 1) + 1] = relu_pack(acc[2], acc[3]);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (ht < 15) {
                a0 = n0;
                a1 = n1;
            }
        }
        // layer 2 in chunks of OC output tiles, then ReLU . w3: step j covers
        // output tiles o0 .. o0 + OC - 1 (o0 = OC * (j / 8)) at k step s = j % 8,
        // which reads chunk column 4s + g of W2P; the next step's fragments
        // are read while this step's MFMAs run
        constexpr int kSteps = 8 * (16 / OC);
        auto frag2 = [&](int j, int oc) {
            const int o0 = OC * (j / 8), s = j % 8;
            return lds8(L1, kW1Bytes + (uint32_t)s * 4 * kChunkCol + 256 * (o0 + oc));
        };
        float y[R];
#pragma unroll
        for (int t = 0; t < R; ++t) y[t] = 0.f;
        f32x4 acc[R][OC];
        bf16x8 a[OC];
#pragma unroll
        for (int oc = 0; oc < OC; ++oc) a[oc] = frag2(0, oc);
#pragma unroll
        for (int j = 0; j < kSteps; ++j) {
            const int o0 = OC * (j / 8), s = j % 8;
            bf16x8 nx[OC];
            if (j + 1 < kSteps) {
#pragma unroll
                for (int oc = 0; oc < OC; ++oc) nx[oc] = frag2(j + 1, oc);
            }
            if (s == 0) {
#pragma unroll
                for (int oc = 0; oc < OC; ++oc) {
                    const f32x4 bias = bcast(vb2, o0 + oc);
#pragma unroll
                    for (int t = 0; t < R; ++t) acc[t][oc] = bias;
                }
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                bf16x8 hb;
                __builtin_memcpy(&hb, hw[t][s], 16);
#pragma unroll
                for (int oc = 0; oc < OC; ++oc) acc[t][oc] = mfma(a[oc], hb, acc[t][oc]);
            }
            if (s == 7) {
#pragma unroll
                for (int oc = 0; oc < OC; ++oc) {
                    const f32x4 w3 = bcast(vw3, o0 + oc);
#pragma unroll
                    for (int t = 0; t < R; ++t)
#pragma unroll
                        for (int e = 0; e < 4; ++e) y[t] += relu(acc[t][oc][e]) * w3[e];
                }
                // keep the sums here (not sunk to their use after the loop,
                // which would hold every accumulator until then)
#pragma unroll
                for (int t = 0; t < R; ++t) asm volatile("" : "+v"(y[t]));
            }
            __builtin_amdgcn_sched_barrier(0);
            if (j + 1 < kSteps) {
#pragma unroll
                for (int oc = 0; oc < OC; ++oc) a[oc] = nx[oc];
            }
        }
        // rows are lane % 16: sum the four lane groups
#pragma unroll
        for (int t = 0; t < R; ++t) {
            float v = y[t];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            const uint64_t row = row0 + 16 * t + r;
            if (g == 0 && row < n) logits[row] = v + b3;
        }
    }
}

// Backward through the activations of a batch: weighted logistic loss on
// the logits, and the bf16 matrices the gradient reductions need, all
// row-major [n][256]: H1 and H2 (post-ReLU), dH2 and dH1 (gradients before the
// ReLU of their layer); dy [n] fp32. The summed weighted loss and db3 are
// reduced per wave and added atomically.

__device__ __forceinline__ void store4(__bf16* p, const f32x4& v) {
    __bf16 t[4] = {(__bf16)v[0], (__bf16)v[1], (__bf16)v[2], (__bf16)v[3]};
    __builtin_memcpy(p, t, 8);
}

__global__ __launch_bounds__(256) void mlp_backward_act_kernel(MlpWeights w, const __bf16* X,
                                                               const float* labels, uint64_t n,
                                                               float pos_weight, float inv_n,
                                                               MlpGrads gr, __bf16* H1, __bf16* H2,
                                                               __bf16* dH2, __bf16* dH1, float* dyout) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    float acc_db3 = 0.f, acc_loss = 0.f;
    Tile T;
    for (uint64_t tile = wave; tile * 16 < n; tile += nwaves) {
        const uint64_t row0 = tile * 16, row = row0 + r;
        const bool valid = row < n;
        const float y = forward_tile<true>(w, X, row0, n, lane, &T);
        float dy = 0.f;
        if (valid) {
            const float lab = labels[row];
            const float p = 1.f / (1.f + __expf(-y));
            const float wt = lab > 0.5f ? pos_weight : 1.f;
            dy = wt * (p - lab) * inv_n;
            const float sp = y > 0.f ? y + log1pf(__expf(-y)) : log1pf(__expf(y));
            if (g == 0) {
                acc_loss += wt * (sp - lab * y);
                acc_db3 += dy;
                dyout[row] = dy;
            }
        }
        f32x4 d2[16];
#pragma unroll
        for (int ot = 0; ot < 16; ++ot) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int u = 16 * ot + 4 * g + t;
                d2[ot][t] = T.h2[ot][t] > 0.f ? dy * w.w3[u] : 0.f;
            }
            if (valid) {
                const uint64_t off = row * kHid + 16 * ot + 4 * g;
                store4(dH2 + off, d2[ot]);
                store4(H2 + off, T.h2[ot]);
                store4(H1 + off, T.h1[ot]);
            }
        }
        // dH1^T = W2 . dH2^T (A = W2Q, B = dH2 in the permuted k order), then ReLU'
#pragma unroll
        for (int ht = 0; ht < 16; ++ht) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                bf16x8 db;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    db[t] = (__bf16)d2[2 * s][t];
                    db[4 + t] = (__bf16)d2[2 * s + 1][t];
                }
                acc = mfma(load8(w.W2Q + (uint64_t)(16 * ht + r) * kHid + 32 * s + 8 * g), db, acc);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = T.h1[ht][t] > 0.f ? acc[t] : 0.f;
            if (valid) store4(dH1 + row * kHid + 16 * ht + 4 * g, acc);
        }
    }
    float b3 = acc_db3, ls = acc_loss;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        b3 += __shfl_xor(b3, d);
        ls += __shfl_xor(ls, d);
    }
    if (lane == 0) {
        atomicAdd(gr.db3, b3);
        atomicAdd(gr.loss, ls);
    }
}

// out[u] += sum over rows of wgt[row] * M[row][u] (wgt == null: weight 1),
// M bf16 [n][256]: one thread per column, rows split across blocks (enough
// blocks to occupy the device; one atomic add per block and column).
__global__ __launch_bounds__(256) void colsum_kernel(const __bf16* M, const float* wgt, uint64_t n,
                                                     uint64_t rows_per_block, float* out) {
    const int u = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_block;
    const uint64_t r1 = r0 + rows_per_block < n ? r0 + rows_per_block : n;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    uint64_t row = r0;
    for (; row + 4 <= r1; row += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float v = (float)M[(row + k) * kHid + u];
            acc[k] += wgt ? wgt[row + k] * v : v;
        }
    }
    for (; row < r1; ++row) {
        const float v = (float)M[row * kHid + u];
        acc[0] += wgt ? wgt[row] * v : v;
    }
    atomicAdd(out + u, (acc[0] + acc[1]) + (acc[2] + acc[3]));
}

// C[a][b] += sum_k A[k][a] * B[k][b] over k in this block's chunk (A, B bf16
// row-major with leading dimensions lda, ldb; C fp32 [a][b], ldc). Block tile
// 64 x 64, 4 waves each 32 x 32 (2 x 2 MFMA tiles), k step 32 through LDS.
__global__ __launch_bounds__(256) void gemm_tn_kernel(const __bf16* A, int lda, const __bf16* B, int ldb,
                                                      float* C, int ldc, uint64_t K, uint64_t kchunk) {
    __shared__ __bf16 As[32][64 + 8];
    __shared__ __bf16 Bs[32][64 + 8];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int a0 = blockIdx.x * 64, b0 = blockIdx.y * 64;
    const uint64_t k0 = (uint64_t)blockIdx.z * kchunk;
    const uint64_t k1 = k0 + kchunk < K ? k0 + kchunk : K;
    const int wa = (wv >> 1) * 32, wb = (wv & 1) * 32;
    f32x4 acc[2][2] = {};
    for (uint64_t kk = k0; kk < k1; kk += 32) {
        // stage 32 x 64 of A and B (each thread: 8 bf16 of one row)
        {
            const int row = tid >> 3, col = (tid & 7) * 8;
            const uint64_t k = kk + row;
            bf16x8 va = {}, vb = {};
            if (k < k1) {
                va = load8(A + k * lda + a0 + col);
                vb = load8(B + k * ldb + b0 + col);
            }
            __builtin_memcpy(&As[row][col], &va, 16);
            __builtin_memcpy(&Bs[row][col], &vb, 16);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                // A operand: row (a) = wa + 16i + r, k = 8g .. +7 ; B: column b = wb + 16j + r
                bf16x8 fa, fb;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    fa[t] = As[8 * g + t][wa + 16 * i + r];
                    fb[t] = Bs[8 * g + t][wb + 16 * j + r];
                }
                acc[i][j] = mfma(fa, fb, acc[i][j]);
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int a = a0 + wa + 16 * i + 4 * g + t, b = b0 + wb + 16 * j + r;
                atomicAdd(C + (uint64_t)a * ldc + b, acc[i][j][t]);
            }
}

// Adam update of fp32 master parameters; grads are zeroed afterwards.
__global__ void adam_kernel(float* p, float* grad, float* m, float* v, uint32_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gd = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gd;
    const float vi = b2 * v[i] + (1.f - b2) * gd * gd;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.f;
}

// bf16 operand layouts from the fp32 master weights W1 [64][256], W2 [256][256].
__global__ void pack_weights_kernel(const float* W1, const float* W2, __bf16* W1T, __bf16* W2P,
                                    __bf16* W2Q) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kHid * kIn) {
        const int h = i / kIn, c = i % kIn;
        W1T[i] = (__bf16)W1[c * kHid + h];
    }
    if (i < kHid * kHid) {
        const int row = i / kHid, col = i % kHid;
        const int s = col / 32, g = (col % 32) / 8, t = col % 8;
        const int u = perm_unit(s, g, t);
        W2P[i] = (__bf16)W2[u * kHid + row];
        W2Q[i] = (__bf16)W2[row * kHid + u];
    }
}

__global__ void f32_to_bf16_kernel(const float* src, __bf16* dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (__bf16)src[i];
}

// ---- launchers ----

// Forward shape: 4 row tiles per weight read (64 rows per group), output
// tiles in pairs, 8 waves per block (2 per SIMD at 244 VGPRs).
constexpr int kFwdR = 4, kFwdOC = 2, kFwdWaves = 8;

hipError_t mlp_launch_forward(const MlpWeights& w, const __bf16* X, uint64_t n, float* logits,
                              int cus, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // one block per CU (the weights fill its LDS), no more blocks than groups
    const uint64_t rows_per_block = 16 * kFwdR * kFwdWaves;
    const uint64_t blocks = (n + rows_per_block - 1) / rows_per_block;
    const int g = (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)cus));
    SABER_LAUNCH((mlp_forward_kernel<kFwdR, kFwdOC, kFwdWaves>), dim3(g), dim3(64 * kFwdWaves), 0, s, w, X,
                       n, logits);
    return hipGetLastError();
}

hipError_t mlp_launch_backward_act(const MlpWeights& w, const __bf16* X, const float* labels,
                                   uint64_t n, float pos_weight, float inv_n, const MlpGrads& gr,
                                   __bf16* H1, __bf16* H2, __bf16* dH2, __bf16* dH1, float* dy,
                                   int grid, hipStream_t s) {
    if (n == 0) return hipSuccess;
    SABER_LAUNCH(mlp_backward_act_kernel, dim3(grid), dim3(256), 0, s, w, X, labels, n,
                       pos_weight, inv_n, gr, H1, H2, dH2, dH1, dy);
    return hipGetLastError();
}

hipError_t mlp_launch_colsum(const __bf16* M, const float* wgt, uint64_t n, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // at most 4096 blocks (same-address atomics per column), at least 64 rows each
    const uint64_t rpb = std::max<uint64_t>(64, (n + 4095) / 4096);
    SABER_LAUNCH(colsum_kernel, dim3((uint32_t)((n + rpb - 1) / rpb)), dim3(256), 0, s, M, wgt, n,
                       rpb, out);
    return hipGetLastError();
}

hipError_t mlp_launch_gemm_tn(const __bf16* A, int lda, int a_dim, const __bf16* B, int ldb, int b_dim,
                              float* C, uint64_t K, int kchunks, hipStream_t s) {
    if (K == 0) return hipSuccess;
    uint64_t kchunk = (K + kchunks - 1) / kchunks;
    kchunk = (kchunk + 31) / 32 * 32;
    const uint32_t nz = (uint32_t)((K + kchunk - 1) / kchunk);
    SABER_LAUNCH(gemm_tn_kernel, dim3(a_dim / 64, b_dim / 64, nz), dim3(256), 0, s, A, lda, B, ldb,
                       C, b_dim, K, kchunk);
    return hipGetLastError();
}

hipError_t mlp_launch_adam(float* p, float* grad, float* m, float* v, uint32_t n, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s) {
    SABER_LAUNCH(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, grad, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_launch_pack(const float* W1, const float* W2, __bf16* W1T, __bf16* W2P, __bf16* W2Q,
                           hipStream_t s) {
    SABER_LAUNCH(pack_weights_kernel, dim3(kHid * kHid / 256), dim3(256), 0, s, W1, W2, W1T, W2P, W2Q);
    return hipGetLastError();
}

hipError_t mlp_launch_to_bf16(const float* src, __bf16* dst, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    SABER_LAUNCH(f32_to_bf16_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace saber
