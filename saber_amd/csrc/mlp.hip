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

// Inference: weights stay in registers for the whole launch, split across
// the four waves of a block: wave v owns hidden units [64v, 64v + 64) of both
// layers (W1T rows: 8 fragments, W2P rows: 32 fragments). Per 16-row tile,
// every wave computes its quarter of layer 1 from the tile's features and
// writes it (bf16, after ReLU) to LDS; after a barrier each wave reads the
// whole layer-1 tile back as layer-2 operands (in the permuted k order W2P
// expects), computes its quarter of layer 2 and its part of the output dot
// product, and wave 0 adds the four parts. LDS is double-buffered by tile.
constexpr int kHs = kHid + 8;  // LDS row stride (bf16) of the layer-1 tile

__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpWeights w, const __bf16* X, uint64_t n,
                                                          float* logits) {
    __shared__ __bf16 Hs[2][16 * kHs];
    __shared__ float part[2][4][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    bf16x8 a1[4][2], a2[4][8];
#pragma unroll
    for (int ht = 0; ht < 4; ++ht)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            a1[ht][ks] = load8(w.W1T + (uint64_t)(64 * wv + 16 * ht + r) * kIn + 32 * ks + 8 * g);
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int s = 0; s < 8; ++s)
            a2[ot][s] = load8(w.W2P + (uint64_t)(64 * wv + 16 * ot + r) * kHid + 32 * s + 8 * g);
    float b1v[4][4], b2v[4][4], w3v[4][4];
#pragma unroll
    for (int t4 = 0; t4 < 4; ++t4)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int u = 64 * wv + 16 * t4 + 4 * g + t;
            b1v[t4][t] = w.b1[u];
            b2v[t4][t] = w.b2[u];
            w3v[t4][t] = w.w3[u];
        }
    const float b3 = *w.b3;
    const uint64_t ntiles = (n + 15) / 16;
    int buf = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x, buf ^= 1) {
        const uint64_t row = tile * 16 + r;
        bf16x8 xb[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) xb[ks] = row < n ? load8(X + row * kIn + 32 * ks + 8 * g) : bf16x8{};
        // layer 1, this wave's 64 units -> LDS
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(a1[ht][0], xb[0], acc);
            acc = mfma(a1[ht][1], xb[1], acc);
            __bf16 hv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float v = acc[t] + b1v[ht][t];
                hv[t] = (__bf16)(v > 0.f ? v : 0.f);
            }
            __builtin_memcpy(&Hs[buf][r * kHs + 64 * wv + 16 * ht + 4 * g], hv, 8);
        }
        __syncthreads();
        // layer 2, this wave's 64 units, operands from the whole layer-1 tile
        f32x4 acc2[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc2[ot] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            bf16x8 hb;
            __builtin_memcpy(&hb, &Hs[buf][r * kHs + 32 * s + 4 * g], 8);
            __builtin_memcpy(reinterpret_cast<char*>(&hb) + 8, &Hs[buf][r * kHs + 32 * s + 16 + 4 * g], 8);
#pragma unroll
            for (int ot = 0; ot < 4; ++ot) acc2[ot] = mfma(a2[ot][s], hb, acc2[ot]);
        }
        float y = 0.f;
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float v = acc2[ot][t] + b2v[ot][t];
                y += (v > 0.f ? v : 0.f) * w3v[ot][t];
            }
        y += __shfl_xor(y, 16);
        y += __shfl_xor(y, 32);
        if (lane < 16) part[buf][wv][lane] = y;
        __syncthreads();
        if (wv == 0 && lane < 16 && row < n)
            logits[row] = part[buf][0][lane] + part[buf][1][lane] + part[buf][2][lane] + part[buf][3][lane] + b3;
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
// M bf16 [n][256]: one thread per column, rows split across blocks.
__global__ __launch_bounds__(256) void colsum_kernel(const __bf16* M, const float* wgt, uint64_t n,
                                                     uint64_t rows_per_block, float* out) {
    const int u = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_block;
    const uint64_t r1 = r0 + rows_per_block < n ? r0 + rows_per_block : n;
    float acc = 0.f;
    for (uint64_t row = r0; row < r1; ++row) {
        const float v = (float)M[row * kHid + u];
        acc += wgt ? wgt[row] * v : v;
    }
    atomicAdd(out + u, acc);
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

hipError_t mlp_launch_forward(const MlpWeights& w, const __bf16* X, uint64_t n, float* logits,
                              int grid, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // one 16-row tile per block per step; grid = blocks resident at once
    const uint64_t tiles = (n + 15) / 16;
    const int g = (int)std::min<uint64_t>(tiles, (uint64_t)grid);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(g), dim3(256), 0, s, w, X, n, logits);
    return hipGetLastError();
}

hipError_t mlp_launch_backward_act(const MlpWeights& w, const __bf16* X, const float* labels,
                                   uint64_t n, float pos_weight, float inv_n, const MlpGrads& gr,
                                   __bf16* H1, __bf16* H2, __bf16* dH2, __bf16* dH1, float* dy,
                                   int grid, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_backward_act_kernel, dim3(grid), dim3(256), 0, s, w, X, labels, n,
                       pos_weight, inv_n, gr, H1, H2, dH2, dH1, dy);
    return hipGetLastError();
}

hipError_t mlp_launch_colsum(const __bf16* M, const float* wgt, uint64_t n, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t rpb = 2048;
    hipLaunchKernelGGL(colsum_kernel, dim3((uint32_t)((n + rpb - 1) / rpb)), dim3(256), 0, s, M, wgt, n,
                       rpb, out);
    return hipGetLastError();
}

hipError_t mlp_launch_gemm_tn(const __bf16* A, int lda, int a_dim, const __bf16* B, int ldb, int b_dim,
                              float* C, uint64_t K, int kchunks, hipStream_t s) {
    if (K == 0) return hipSuccess;
    uint64_t kchunk = (K + kchunks - 1) / kchunks;
    kchunk = (kchunk + 31) / 32 * 32;
    const uint32_t nz = (uint32_t)((K + kchunk - 1) / kchunk);
    hipLaunchKernelGGL(gemm_tn_kernel, dim3(a_dim / 64, b_dim / 64, nz), dim3(256), 0, s, A, lda, B, ldb,
                       C, b_dim, K, kchunk);
    return hipGetLastError();
}

hipError_t mlp_launch_adam(float* p, float* grad, float* m, float* v, uint32_t n, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, grad, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_launch_pack(const float* W1, const float* W2, __bf16* W1T, __bf16* W2P, __bf16* W2Q,
                           hipStream_t s) {
    hipLaunchKernelGGL(pack_weights_kernel, dim3(kHid * kHid / 256), dim3(256), 0, s, W1, W2, W1T, W2P, W2Q);
    return hipGetLastError();
}

hipError_t mlp_launch_to_bf16(const float* src, __bf16* dst, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace saber
