// Message-passing network for fixpoint warm starts: FP32 forward, backward and Adam,
// all in hand-written kernels.
//
// Nodes of a batch of grids are stacked (n rows); the graph is a row-normalised
// weighted adjacency (a_fwd[s] = K_s / sum of the row's K) in CSR form, plus its
// transpose weights a_bwd for the backward pass.  Layers (hidden width 64):
//   H0 = relu(X W_in + b_in)                      X: n x F features
//   Y  = H_{l-1} [W_l | U_l]                       (n x 128, one GEMM)
//   H_l = relu(A Y[:, :64] + Y[:, 64:] + b_l)      l = 1..4
//   out = H_4 w_out + b_out                        (n)
// Loss: mean squared error against the target.  Weight gradients are reduced in
// two stages (per-block partials, then a fixed-order sum), so training is
// deterministic.
//
// Parameter vector: W_in[F*64], b_in[64], {W_l|U_l [64*128], b_l[64]} x 4, w_out[64], b_out[1].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "nd_internal.h"

namespace {

constexpr int kH = 64;          // hidden width
constexpr int kLayers = 4;
constexpr int kWave = 64;

// ---- dense: C[M x N] = A[M x K] * op(B) (+ bias, relu) ---------------------------------
// op(B) = B[K x N] (TRANSB = false) or B[N x K]^T (TRANSB = true).  64x64 tiles, 256
// threads, 4x4 outputs per thread, K in steps of 16 through LDS.
template <bool TRANSB>
__global__ void __launch_bounds__(256) k_sgemm(int64_t M, int N, int K, const float* __restrict__ A, int lda,
                                               const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                               const float* __restrict__ bias, int relu) {
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int col0 = blockIdx.y * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int i = threadIdx.x; i < 16 * 64; i += 256) {
            const int kk = i % 16, rr = i / 16;      // A tile: 64 rows x 16 k
            const int64_t r = row0 + rr;
            As[kk][rr] = (r < M && k0 + kk < K) ? A[r * lda + k0 + kk] : 0.0f;
            const int cc = i % 64, kb = i / 64;      // B tile: 16 k x 64 cols
            float bv = 0.0f;
            if (col0 + cc < N && k0 + kb < K)
                bv = TRANSB ? B[static_cast<int64_t>(col0 + cc) * ldb + k0 + kb] : B[static_cast<int64_t>(k0 + kb) * ldb + col0 + cc];
            Bs[kb][cc] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t r = row0 + ty * 4 + i;
        if (r >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + tx * 4 + j;
            if (c >= N) continue;
            float v = acc[i][j] + (bias ? bias[c] : 0.0f);
            if (relu) v = fmaxf(v, 0.0f);
            C[r * ldc + c] = v;
        }
    }
}

// ---- MFMA (v_mfma_f32_16x16x4f32) dense products ---------------------------------------
// Operand layout of one 16x16x4 step in a 64-lane wave: lane l holds A[i][k] and B[k][j]
// with i = j = l % 16, k = l / 16; the 4 accumulators of lane l are D[4*(l/16) + r][l % 16].
typedef float floatx4 __attribute__((ext_vector_type(4)));

// C[M x N] = A[M x K] * op(B) (+ bias, relu), N = 64 or 128, K <= 128.  op(B) = B[K x N]
// (TRANSB false) or B[N x K]^T.  256 threads: a workgroup owns 128 rows, each wave 32 rows
// (2 row tiles) x N columns; A's rows and all of op(B) are staged in LDS with strides
// padded so the fragment reads are free of bank conflicts.
// dynamic LDS floats of k_gemm_mfma: A rows (stride Kp + 2) and op(B) (stride N + 16)
inline size_t gemm_mfma_lds_floats(int K, int N) {
    const int Kp = (K + 3) & ~3;
    return static_cast<size_t>(128) * (Kp + 2) + static_cast<size_t>(Kp) * (N + 16);
}

template <bool TRANSB, int N>
__global__ void __launch_bounds__(256) k_gemm_mfma(int64_t M, int K, const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                                   const float* __restrict__ bias, int relu) {
    constexpr int SB = N + 16, CT = N / 16;
    extern __shared__ float smem_f[];
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    const int Kp = (K + 3) & ~3;
    const int SA = Kp + 2;
    float* As = smem_f;
    float* Bs = smem_f + 128 * SA;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * 128;
    for (int i = t; i < 128 * Kp; i += 256) {
        const int r = i / Kp, k = i % Kp;
        As[r * SA + k] = (row0 + r < M && k < K) ? A[(row0 + r) * lda + k] : 0.0f;
    }
    for (int i = t; i < Kp * N; i += 256) {
        const int k = i / N, j = i % N;
        float v = 0.0f;
        if (k < K) v = TRANSB ? B[static_cast<int64_t>(j) * ldb + k] : B[static_cast<int64_t>(k) * ldb + j];
        Bs[k * SB + j] = v;
    }
    __syncthreads();
    floatx4 acc[2][CT];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int li = lane % 16, lk = lane / 16;
    for (int k0 = 0; k0 < Kp; k0 += 4) {
        float a[2], b[CT];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) a[rt] = As[(w * 32 + rt * 16 + li) * SA + k0 + lk];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) b[ct] = Bs[(k0 + lk) * SB + ct * 16 + li];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
    }
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = row0 + w * 32 + rt * 16 + 4 * lk + r;
                const int col = ct * 16 + li;
                if (row < M) {
                    float v = acc[rt][ct][r] + (bias ? bias[col] : 0.0f);
                    if (relu) v = fmaxf(v, 0.0f);
                    C[row * ldc + col] = v;
                }
            }
}

// C[M x N] = A[M x K] * op(B) (+ bias, relu) for the network's full-width products:
// K = 64, N = 128 (a hidden layer's [W | U]) and K = 128, N = 64 (the backward's
// [dYa | g] [W | U]^T), A dense with K columns.  op(B) is staged once per workgroup in
// LDS; A goes straight from memory into the MFMA operand registers: lane (li, lk) of a
// wave loads row li, columns K/4*lk .. K/4*lk + K/4 - 1 as 16-byte loads, and step s of
// the 16x16x4 products pairs column K/4*lk + s with row K/4*lk + s of op(B), so every
// k enters once.  Each wave takes 16-row tiles, grid-stride (the grid is sized to the
// chip, not to M).  B's odd stride N + 1 spreads the four k rows of one step over the
// LDS banks.
template <int K, int N, bool TRANSB>
__global__ void __launch_bounds__(256) k_gemm_reg(int64_t M, const float* __restrict__ A, const float* __restrict__ B,
                                                  int ldb, float* __restrict__ C, int ldc, const float* __restrict__ bias,
                                                  int relu) {
    constexpr int SB = N + 1, CT = N / 16, KS = K / 4;
    __shared__ float Bs[K * SB];
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    for (int i = t; i < K * N; i += 256) {
        const int k = i / N, j = i % N;
        Bs[k * SB + j] = TRANSB ? B[static_cast<int64_t>(j) * ldb + k] : B[static_cast<int64_t>(k) * ldb + j];
    }
    __syncthreads();
    const int li = lane % 16, lk = lane / 16;
    const int64_t ntiles = (M + 15) / 16;
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + w; tile < ntiles; tile += static_cast<int64_t>(gridDim.x) * 4) {
        const int64_t row = tile * 16 + li;
        float a[KS];
        if (row < M) {
            const float4* ap = reinterpret_cast<const float4*>(A + row * K + KS * lk);
#pragma unroll
            for (int q = 0; q < KS / 4; ++q) {
                const float4 v = ap[q];
                a[4 * q] = v.x;
                a[4 * q + 1] = v.y;
                a[4 * q + 2] = v.z;
                a[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < KS; ++q) a[q] = 0.0f;
        }
        floatx4 acc[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        // four steps at a time: the fence keeps the compiler from hoisting every B read of
        // the tile into registers at once (which halves the waves a SIMD can hold)
#pragma unroll
        for (int s0 = 0; s0 < KS; s0 += 4) {
#pragma unroll
            for (int s = s0; s < s0 + 4; ++s)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], Bs[(KS * lk + s) * SB + ct * 16 + li], acc[ct], 0,
                                                                   0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t ro = tile * 16 + 4 * lk + r;
                const int col = ct * 16 + li;
                if (ro < M) {
                    float v = acc[ct][r] + (bias ? bias[col] : 0.0f);
                    if (relu) v = fmaxf(v, 0.0f);
                    C[ro * ldc + col] = v;
                }
            }
    }
}

// partial[b] = A[rows_b]^T * B[rows_b] on MFMA: KA <= 64 output rows (one 16-row tile per
// wave), NB <= 128 output columns, part_rows(M) rows of the product's inner dimension
// per workgroup, staged through LDS 64 rows at a time.
constexpr int kTnChunk = 64;      // rows staged per LDS round
template <int NBT>
__global__ void __launch_bounds__(256) k_gemm_tn_mfma(int64_t M, int KA, int NB, const float* __restrict__ A, int lda,
                                                      const float* __restrict__ B, int ldb, float* __restrict__ part,
                                                      int rows_per_block) {
    constexpr int SA = 64 + 16, SB = NBT * 16 + 16;
    __shared__ float As[kTnChunk * SA];
    __shared__ float Bs[kTnChunk * SB];
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    const int li = lane % 16, lk = lane / 16;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < M ? r0 + rows_per_block : M;
    floatx4 acc[NBT];
#pragma unroll
    for (int ct = 0; ct < NBT; ++ct) acc[ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t rb = r0; rb < r1; rb += kTnChunk) {
        for (int i = t; i < kTnChunk * 64; i += 256) {
            const int rr = i / 64, c = i % 64;
            As[rr * SA + c] = (rb + rr < r1 && c < KA) ? A[(rb + rr) * lda + c] : 0.0f;
        }
        for (int i = t; i < kTnChunk * NBT * 16; i += 256) {
            const int rr = i / (NBT * 16), c = i % (NBT * 16);
            Bs[rr * SB + c] = (rb + rr < r1 && c < NB) ? B[(rb + rr) * ldb + c] : 0.0f;
        }
        __syncthreads();
        // output rows w*16 .. w*16+15 of A^T B: A^T[i][k] = As[k][i]
        for (int k0 = 0; k0 < kTnChunk; k0 += 4) {
            const float a = As[(k0 + lk) * SA + w * 16 + li];
#pragma unroll
            for (int ct = 0; ct < NBT; ++ct)
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bs[(k0 + lk) * SB + ct * 16 + li], acc[ct], 0, 0, 0);
        }
        __syncthreads();
    }
    const int nout = KA * NB;
#pragma unroll
    for (int ct = 0; ct < NBT; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = w * 16 + 4 * lk + r, j = ct * 16 + li;
            if (i < KA && j < NB) part[static_cast<int64_t>(blockIdx.x) * nout + i * NB + j] = acc[ct][r];
        }
}

// ---- dense: partial[b] = A[rows_b]^T * B[rows_b]  (KA x NB, KA*NB <= 8192) ---------------
// Rows of the inner dimension per workgroup of these two-stage sums (A^T B, column and
// vector sums): enough workgroups to fill the chip, at most kMaxPartBlocks of them, at
// least 64 rows each.  The partials are then added in a fixed order (k_reduce_partials),
// so every sum is deterministic for a given M.
constexpr int kMaxPartBlocks = 1024;
inline int part_rows(int64_t M) {
    const int64_t r = (M + kMaxPartBlocks - 1) / kMaxPartBlocks;
    return static_cast<int>(std::max<int64_t>(64, (r + 63) / 64 * 64));
}
__global__ void __launch_bounds__(256) k_sgemm_tn_partial(int64_t M, int KA, int NB, const float* __restrict__ A, int lda,
                                                          const float* __restrict__ B, int ldb, float* __restrict__ part,
                                                          int rows_per_block) {
    __shared__ float As[16][128];
    __shared__ float Bs[16][128];
    const int nout = KA * NB;
    float acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.0f;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r1 = std::min<int64_t>(r0 + rows_per_block, M);
    for (int64_t rb = r0; rb < r1; rb += 16) {
        for (int i = threadIdx.x; i < 16 * 128; i += 256) {
            const int rr = i / 128, c = i % 128;
            const int64_t r = rb + rr;
            As[rr][c] = (r < r1 && c < KA) ? A[r * lda + c] : 0.0f;
            Bs[rr][c] = (r < r1 && c < NB) ? B[r * ldb + c] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int o = threadIdx.x + q * 256;
            if (o < nout) {
                const int i = o / NB, j = o % NB;
                float s = 0.0f;
#pragma unroll
                for (int rr = 0; rr < 16; ++rr) s += As[rr][i] * Bs[rr][j];
                acc[q] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int o = threadIdx.x + q * 256;
        if (o < nout) part[static_cast<int64_t>(blockIdx.x) * nout + o] = acc[q];
    }
}

// out[i] (+)= sum_b part[b][i], in block order
// (64 outputs per workgroup of kRedWaves waves: wave w adds the partials of blocks
// [nb*w/kRedWaves, nb*(w+1)/kRedWaves) in order, then the wave sums are added in wave order)
constexpr int kRedWaves = 16;
__global__ void __launch_bounds__(kRedWaves * 64) k_reduce_partials(int nb, int nout, const float* __restrict__ part,
                                                                    float* __restrict__ out, int add) {
    __shared__ float sh[kRedWaves][kWave];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int i = blockIdx.x * kWave + lane;
    const int b0 = nb * w / kRedWaves, b1 = nb * (w + 1) / kRedWaves;
    float s = 0.0f;
    if (i < nout) {
#pragma unroll 4
        for (int b = b0; b < b1; ++b) s += part[static_cast<int64_t>(b) * nout + i];
    }
    sh[w][lane] = s;
    __syncthreads();
    if (w == 0 && i < nout) {
        float t = sh[0][lane];
        for (int q = 1; q < kRedWaves; ++q) t += sh[q][lane];
        out[i] = add ? out[i] + t : t;
    }
}

// column sums of G[M x 64] into per-block partials (bias gradients)
__global__ void __launch_bounds__(256) k_colsum_partial(int64_t M, const float* __restrict__ G, int ldg, float* __restrict__ part,
                                                        int rows_per_block) {
    __shared__ float sh[4][kH];
    const int c = threadIdx.x % kH, w = threadIdx.x / kH;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r1 = std::min<int64_t>(r0 + rows_per_block, M);
    float s = 0.0f;
    for (int64_t r = r0 + w; r < r1; r += 4) s += G[r * ldg + c];
    sh[w][c] = s;
    __syncthreads();
    if (w == 0) part[static_cast<int64_t>(blockIdx.x) * kH + c] = sh[0][c] + sh[1][c] + sh[2][c] + sh[3][c];
}

// per-block sums of a vector (fixed tree inside the block)
__global__ void __launch_bounds__(256) k_vsum_partial(int64_t n, const float* __restrict__ x, float* __restrict__ part,
                                                      int rows_per_block) {
    __shared__ float sh[256];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r1 = std::min<int64_t>(r0 + rows_per_block, n);
    float s = 0.0f;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) s += x[r];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Workgroups are dispatched to the chip's 8 XCDs round robin (workgroup b to XCD b % 8).
// xcd_block(b, nb) renumbers them so that each XCD works through one contiguous eighth
// of the grid: rows of one grid, which read each other's rows, then share one L2.
// A bijection of [0, nb).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t x = b % 8, q = nb / 8, r = nb % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

// ---- sparse: out[i, :] = sum_s a[s] * Y[nbr[s], 0:64]  (+ Y[i, 64:128] + bias, relu) -------
// 16 lanes per node, four features per lane (16-byte loads), so a wave works on four
// nodes at once; each lane sums its features over the node's entries in order, with four
// entries' rows in flight.  Y, out and bias must be 16-byte aligned (ldy, ldo multiples of 4).
__global__ void __launch_bounds__(256) k_spmm(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ nbr,
                                              const float* __restrict__ a, const float* __restrict__ Y, int ldy,
                                              float* __restrict__ out, int ldo, const float* __restrict__ bias, int fused) {
    const int64_t i = xcd_block(blockIdx.x, gridDim.x) * 16 + threadIdx.x / 16;
    const int c = static_cast<int>(threadIdx.x % 16) * 4;
    if (i >= n) return;
    auto row = [&](int32_t j) { return *reinterpret_cast<const float4*>(Y + static_cast<int64_t>(j) * ldy + c); };
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto add = [&](float w, const float4& y) {
        s.x += w * y.x;
        s.y += w * y.y;
        s.z += w * y.z;
        s.w += w * y.w;
    };
    int32_t k = rp[i];
    const int32_t k1 = rp[i + 1];
    for (; k + 4 <= k1; k += 4) {
        const int32_t j0 = nbr[k], j1 = nbr[k + 1], j2 = nbr[k + 2], j3 = nbr[k + 3];
        const float a0 = a[k], a1 = a[k + 1], a2 = a[k + 2], a3 = a[k + 3];
        const float4 y0 = row(j0), y1 = row(j1), y2 = row(j2), y3 = row(j3);
        add(a0, y0);
        add(a1, y1);
        add(a2, y2);
        add(a3, y3);
    }
    for (; k < k1; ++k) add(a[k], row(nbr[k]));
    if (fused) {
        const float4 y = *reinterpret_cast<const float4*>(Y + i * ldy + kH + c);
        const float4 b = *reinterpret_cast<const float4*>(bias + c);
        s.x = fmaxf(s.x + (y.x + b.x), 0.0f);
        s.y = fmaxf(s.y + (y.y + b.y), 0.0f);
        s.z = fmaxf(s.z + (y.z + b.z), 0.0f);
        s.w = fmaxf(s.w + (y.w + b.w), 0.0f);
    }
    *reinterpret_cast<float4*>(out + i * ldo + c) = s;
}

// g = dH * relu'(H): G[i, c] = H[i, c] > 0 ? DH[i, c] : 0, also copied to DY[i, 64 + c] when DY is given
__global__ void k_relu_grad(int64_t n, const float* __restrict__ DH, const float* __restrict__ H, float* __restrict__ G,
                            float* __restrict__ DY) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * kH) return;
    const int64_t i = t / kH;
    const int c = static_cast<int>(t % kH);
    const float g = H[t] > 0.0f ? DH[t] : 0.0f;
    G[t] = g;
    if (DY) DY[i * 2 * kH + kH + c] = g;
}

// out[i] = H[i, :] . w + b ; optional MSE gradient against target and loss partials
__global__ void __launch_bounds__(256) k_head(int64_t n, const float* __restrict__ H, const float* __restrict__ w,
                                              const float* __restrict__ b, float* __restrict__ out,
                                              const float* __restrict__ target, float* __restrict__ dout,
                                              float* __restrict__ dH, float* __restrict__ loss_part) {
    __shared__ float sl[4];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / kWave;
    const int c = threadIdx.x % kWave;
    float v = 0.0f, l = 0.0f;
    if (i < n) {
        v = H[i * kH + c] * w[c];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        v += b[0];
        if (c == 0 && out) out[i] = v;
        if (target) {
            const float e = v - target[i];
            const float g = 2.0f * e / static_cast<float>(n);
            if (c == 0) { dout[i] = g; l = e * e; }
            dH[i * kH + c] = g * w[c];
        }
    }
    if (loss_part) {
        if (c == 0) sl[threadIdx.x / kWave] = l;
        __syncthreads();
        if (threadIdx.x == 0) loss_part[blockIdx.x] = sl[0] + sl[1] + sl[2] + sl[3];
    }
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned nblk(int64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

// ---- per-node inputs of the network (nd_ws_features) ------------------------------------
// One thread per node, entries in CSR order: D = theta[j] - theta[i]; s0 = sum w (sin D - D),
// sum w D^2, max |D|, kdeg = sum w; the columns of X that need only the node's own entries.
constexpr int kFeat = 8;
__global__ void __launch_bounds__(256) k_ws_features(int64_t n, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ nbr, const double* __restrict__ w,
                                                     const double* __restrict__ th, const double* __restrict__ P,
                                                     const double* __restrict__ gen, double* __restrict__ s0,
                                                     double* __restrict__ X) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double ti = th[i];
    double s = 0.0, sq = 0.0, mx = 0.0, kd = 0.0;
    const int32_t k1 = rp[i + 1];
    for (int32_t k = rp[i]; k < k1; ++k) {
        const double D = th[nbr[k]] - ti, wk = w[k];
        s += wk * (sin(D) - D);
        sq += wk * D * D;
        mx = fmax(mx, fabs(D));
        kd += wk;
    }
    s0[i] = s;
    double* x = X + i * kFeat;
    x[0] = P[i];
    x[1] = gen[i];
    x[2] = kd > 0.0 ? log(kd) : 0.0;        // a node without entries: 0
    x[3] = static_cast<double>(k1 - rp[i]) / 6.0;
    x[4] = s;
    x[6] = sq;
    x[7] = mx;
}

// |F_i| / scale_i of the fixpoint residual F_i = P_i + sum_s w_s sin(theta[nbr_s] - theta_i),
// the quantity the batched Newton solver drives to its tolerance (entries in CSR order)
__global__ void __launch_bounds__(256) k_ws_residual(int64_t n, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ nbr, const double* __restrict__ w,
                                                     const double* __restrict__ th, const double* __restrict__ P,
                                                     const double* __restrict__ scale, double* __restrict__ r) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double ti = th[i];
    double F = P[i];
    const int32_t k1 = rp[i + 1];
    for (int32_t k = rp[i]; k < k1; ++k) F += w[k] * sin(th[nbr[k]] - ti);
    r[i] = fabs(F) / scale[i];
}

// the K-weighted neighbour mean of s0 (X column 5), after k_ws_features
__global__ void __launch_bounds__(256) k_ws_features_nbr(int64_t n, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ nbr, const double* __restrict__ w,
                                                         const double* __restrict__ s0, double* __restrict__ X) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    double ns = 0.0, kd = 0.0;
    const int32_t k1 = rp[i + 1];
    for (int32_t k = rp[i]; k < k1; ++k) {
        const double wk = w[k];
        ns += wk * s0[nbr[k]];
        kd += wk;
    }
    X[i * kFeat + 5] = kd > 0.0 ? ns / kd : 0.0;
}

struct Offsets {
    int64_t win, bin, wl[kLayers], bl[kLayers], wout, bout, total;
};
Offsets offsets(int F) {
    Offsets o{};
    int64_t k = 0;
    o.win = k; k += static_cast<int64_t>(F) * kH;
    o.bin = k; k += kH;
    for (int l = 0; l < kLayers; ++l) { o.wl[l] = k; k += kH * 2 * kH; o.bl[l] = k; k += kH; }
    o.wout = k; k += kH;
    o.bout = k; k += 1;
    o.total = k;
    return o;
}

int herr(hipError_t e, const char* w) {
    return e == hipSuccess ? ND_OK : nd::set_error(ND_ERR_DEVICE, "%s: %s", w, hipGetErrorString(e));
}

template <bool TRANSB, int N>
void gemm_mfma(hipStream_t s, int64_t M, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
               const float* bias, int relu) {
    const size_t lds = gemm_mfma_lds_floats(K, N) * sizeof(float);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_gemm_mfma<TRANSB, N>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL((k_gemm_mfma<TRANSB, N>), dim3(nblk(M, 128)), dim3(256), lds, s, M, K, A, lda, B, ldb, C, ldc,
                       bias, relu);
}

// C = A op(B): MFMA for the network's shapes (N = 64 or 128 columns, K <= 128), the
// LDS-tiled FP32 kernel otherwise
void gemm(hipStream_t s, bool transb, int64_t M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
          int ldc, const float* bias, int relu) {
    const bool dense16 = lda == K && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    if (dense16 && ((K == 64 && N == 128 && !transb) || (K == 128 && N == 64 && transb))) {
        // 16-row tiles, four per workgroup; four workgroups per CU fit their LDS
        const unsigned nb = static_cast<unsigned>(std::min<int64_t>((M + 63) / 64, 1024));
        if (K == 64)
            hipLaunchKernelGGL((k_gemm_reg<64, 128, false>), dim3(nb), dim3(256), 0, s, M, A, B, ldb, C, ldc, bias, relu);
        else
            hipLaunchKernelGGL((k_gemm_reg<128, 64, true>), dim3(nb), dim3(256), 0, s, M, A, B, ldb, C, ldc, bias, relu);
        return;
    }
    if (K <= 128 && (N == 64 || N == 128)) {
        if (N == 64) {
            if (transb) gemm_mfma<true, 64>(s, M, K, A, lda, B, ldb, C, ldc, bias, relu);
            else gemm_mfma<false, 64>(s, M, K, A, lda, B, ldb, C, ldc, bias, relu);
        } else {
            if (transb) gemm_mfma<true, 128>(s, M, K, A, lda, B, ldb, C, ldc, bias, relu);
            else gemm_mfma<false, 128>(s, M, K, A, lda, B, ldb, C, ldc, bias, relu);
        }
        return;
    }
    dim3 grid(nblk(M, 64), nblk(N, 64));
    if (transb) hipLaunchKernelGGL(k_sgemm<true>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, relu);
    else hipLaunchKernelGGL(k_sgemm<false>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, relu);
}

// out (+)= A^T B over all M rows: per-block partials, then a fixed-order sum (deterministic).
// MFMA for KA <= 64 and 16 <= NB <= 128; the FP32 VALU kernel for the thin cases.
void gemm_tn(hipStream_t s, int64_t M, int KA, int NB, const float* A, int lda, const float* B, int ldb, float* part,
             float* out, int add) {
    const int rows = part_rows(M);
    const unsigned nb = std::max(1u, nblk(M, rows));
    if (KA <= 64 && NB >= 16 && NB <= 128) {
        const int nbt = (NB + 15) / 16;
        if (nbt <= 4)
            hipLaunchKernelGGL(k_gemm_tn_mfma<4>, dim3(nb), dim3(256), 0, s, M, KA, NB, A, lda, B, ldb, part, rows);
        else
            hipLaunchKernelGGL(k_gemm_tn_mfma<8>, dim3(nb), dim3(256), 0, s, M, KA, NB, A, lda, B, ldb, part, rows);
    } else {
        hipLaunchKernelGGL(k_sgemm_tn_partial, dim3(nb), dim3(256), 0, s, M, KA, NB, A, lda, B, ldb, part, rows);
    }
    hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(KA * NB, kWave)), dim3(kRedWaves * 64), 0, s, static_cast<int>(nb),
                       KA * NB, part, out, add);
}

void colsum(hipStream_t s, int64_t M, const float* G, int ldg, float* part, float* out) {
    const int rows = part_rows(M);
    const unsigned nb = std::max(1u, nblk(M, rows));
    hipLaunchKernelGGL(k_colsum_partial, dim3(nb), dim3(256), 0, s, M, G, ldg, part, rows);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedWaves * 64), 0, s, static_cast<int>(nb), kH, part, out, 0);
}

}  // namespace

extern "C" {

int64_t nd_mp_param_count(int32_t nfeat) { return offsets(nfeat).total; }

int64_t nd_mp_workspace_floats(int64_t nnodes) {
    // activations H0..H4 (5*64) + Y (128) per node, backward: g (64), [dYa|g] (128), dH (64),
    // head gradient (1) + partial buffers
    const int64_t nb = std::min<int64_t>(kMaxPartBlocks, (nnodes + 63) / 64);
    return nnodes * (5 * kH + 2 * kH + kH + 2 * kH + kH + 1) + (nb + 2) * (kH * 2 * kH) + nnodes / 4 + 8;
}

int nd_mp_forward(int64_t n, const int32_t* rowptr, const int32_t* nbr, const float* a_fwd, const float* X, int32_t nfeat,
                  const float* params, float* work, float* out, void* stream) {
    if (n < 1 || n > INT32_MAX - 8 || !rowptr || !X || !params || !work || !out || nfeat < 1 || nfeat > 128)
        return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    if ((reinterpret_cast<uintptr_t>(params) | reinterpret_cast<uintptr_t>(work)) & 15)
        return nd::set_error(ND_ERR_ARGUMENT, "params and work must be 16-byte aligned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Offsets o = offsets(nfeat);
    float* H = work;                        // H_l at H + l*n*64
    float* Y = work + 5 * n * kH;           // n x 128
    gemm(s, false, n, kH, nfeat, X, nfeat, params + o.win, kH, H, kH, params + o.bin, 1);
    for (int l = 0; l < kLayers; ++l) {
        gemm(s, false, n, 2 * kH, kH, H + l * n * kH, kH, params + o.wl[l], 2 * kH, Y, 2 * kH, nullptr, 0);
        hipLaunchKernelGGL(k_spmm, dim3(nblk(n, 16)), dim3(256), 0, s, n, rowptr, nbr, a_fwd, Y, 2 * kH,
                           H + (l + 1) * n * kH, kH, params + o.bl[l], 1);
    }
    hipLaunchKernelGGL(k_head, dim3(nblk(n, 4)), dim3(256), 0, s, n, H + 4 * n * kH, params + o.wout, params + o.bout,
                       out, nullptr, nullptr, nullptr, nullptr);
    return herr(hipGetLastError(), "nd_mp_forward");
}

// Gradient of the mean squared error; the forward pass must have filled `work`
// with the same inputs.  loss_part: n/4 + 1 floats of per-block loss sums.
int nd_mp_backward(int64_t n, const int32_t* rowptr, const int32_t* nbr, const float* a_bwd, const float* X,
                   int32_t nfeat, const float* params, const float* target, float* work, float* grads,
                   float* loss_part, void* stream) {
    if (n < 1 || n > INT32_MAX - 8 || !rowptr || !X || !params || !work || !grads || !target || !loss_part || nfeat < 1 ||
        nfeat > 128)
        return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    if ((reinterpret_cast<uintptr_t>(params) | reinterpret_cast<uintptr_t>(work)) & 15)
        return nd::set_error(ND_ERR_ARGUMENT, "params and work must be 16-byte aligned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Offsets o = offsets(nfeat);
    float* H = work;
    float* G = work + 5 * n * kH + n * 2 * kH;   // g: n x 64
    float* DY = G + n * kH;                       // [dYa | g]: n x 128
    float* DH = DY + n * 2 * kH;                  // dH: n x 64
    float* dout = DH + n * kH;                    // n
    float* part = dout + n;                       // partials
    hipLaunchKernelGGL(k_head, dim3(nblk(n, 4)), dim3(256), 0, s, n, H + 4 * n * kH, params + o.wout, params + o.bout,
                       nullptr, target, dout, DH, loss_part);
    gemm_tn(s, n, kH, 1, H + 4 * n * kH, kH, dout, 1, part, grads + o.wout, 0);
    {
        const int rows = part_rows(n);
        const unsigned nb = std::max(1u, nblk(n, rows));
        hipLaunchKernelGGL(k_vsum_partial, dim3(nb), dim3(256), 0, s, n, dout, part, rows);
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedWaves * 64), 0, s, static_cast<int>(nb), 1, part,
                           grads + o.bout, 0);
    }
    for (int l = kLayers - 1; l >= 0; --l) {
        // g = dH_{l+1} * relu'(H_{l+1}), into G and the right half of [dYa | g]
        hipLaunchKernelGGL(k_relu_grad, dim3(nblk(n * kH, 256)), dim3(256), 0, s, n, DH, H + (l + 1) * n * kH, G, DY);
        colsum(s, n, G, kH, part, grads + o.bl[l]);
        // dYa = A^T g into the left half
        hipLaunchKernelGGL(k_spmm, dim3(nblk(n, 16)), dim3(256), 0, s, n, rowptr, nbr, a_bwd, G, kH, DY, 2 * kH,
                           nullptr, 0);
        // d[W|U] = H_l^T [dYa | g];  dH_l = [dYa | g] [W|U]^T
        gemm_tn(s, n, kH, 2 * kH, H + l * n * kH, kH, DY, 2 * kH, part, grads + o.wl[l], 0);
        gemm(s, true, n, kH, 2 * kH, DY, 2 * kH, params + o.wl[l], 2 * kH, DH, kH, nullptr, 0);
    }
    hipLaunchKernelGGL(k_relu_grad, dim3(nblk(n * kH, 256)), dim3(256), 0, s, n, DH, H, G, nullptr);
    colsum(s, n, G, kH, part, grads + o.bin);
    gemm_tn(s, n, nfeat, kH, X, nfeat, G, kH, part, grads + o.win, 0);
    return herr(hipGetLastError(), "nd_mp_backward");
}

int nd_ws_features(int64_t n, const int32_t* rowptr, const int32_t* nbr, const double* w, const double* theta,
                   const double* P, const double* is_gen, double* s0, double* X, void* stream) {
    if (n < 1 || !rowptr || !nbr || !w || !theta || !P || !is_gen || !s0 || !X)
        return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_ws_features, dim3(nblk(n, 256)), dim3(256), 0, s, n, rowptr, nbr, w, theta, P, is_gen, s0, X);
    hipLaunchKernelGGL(k_ws_features_nbr, dim3(nblk(n, 256)), dim3(256), 0, s, n, rowptr, nbr, w, s0, X);
    return herr(hipGetLastError(), "nd_ws_features");
}

int nd_ws_residual(int64_t n, const int32_t* rowptr, const int32_t* nbr, const double* w, const double* theta,
                   const double* P, const double* scale, double* r, void* stream) {
    if (n < 1 || !rowptr || !nbr || !w || !theta || !P || !scale || !r) return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    hipLaunchKernelGGL(k_ws_residual, dim3(nblk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, rowptr, nbr, w,
                       theta, P, scale, r);
    return herr(hipGetLastError(), "nd_ws_residual");
}

int nd_mp_adam(int64_t np, float* params, const float* grads, float* m, float* v, float lr, float b1, float b2,
               float eps, int64_t step, void* stream) {
    if (np < 1 || !params || !grads || !m || !v || step < 1) return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(step));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(np, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), np, params, grads,
                       m, v, lr, b1, b2, eps, c1, c2);
    return herr(hipGetLastError(), "nd_mp_adam");
}

}  // extern "C"
