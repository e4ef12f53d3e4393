// encoder.hip -- K9: the learned initial-pose proposal (Siamese point encoder).
//
// No counterpart in the reference: the network takes the role of the coarse
// alignment that precedes fine registration (Super4PCS, scripts/alignment.sh:107).
// Per point 3 -> 64 -> 128 -> 256 (ReLU), max-pool per cloud, head 512 -> 256 -> 6
// (axis-angle, translation).  Layer 1 and the head run in FP32; layers 2 and 3 run
// on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, from BF16 copies
// of W2/W3 and BF16 activations h1/h2.
//
// Forward (k_enc_points): one block per kRows (128) points of one cloud.  Layer 1
// writes h1 (bf16) to LDS; each of the 4 waves computes a 32-column slab of layer
// 2 for all rows (its B fragments stay in registers) and writes h2 (bf16) to LDS;
// each wave then computes a 64-column slab of layer 3 and folds bias, ReLU and
// the max-pool (with the lowest point index on ties) into registers, so only 256
// (max, argmax) pairs per block leave the chip.  At 128 rows a block holds 53 KB
// of LDS and two or three blocks share a CU, so one block's MFMA work overlaps
// another's epilogues; at 256 rows (106 KB) a CU held one.
//
// Backward: the max-pool sends each channel's gradient to one point, so the
// per-point layers only need the <= 256 argmax rows per cloud (k_enc_rows
// recomputes their activations).  Weight gradients are sums of outer products
// over pairs / rows, done in two deterministic stages (fixed-order partials).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lcr_kernels.hpp"

namespace lcr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 128;      // points per forward block (two blocks' LDS fit a CU)
constexpr int kH1S = 64 + 8;    // LDS row strides (bf16), padded
constexpr int kH2S = 128 + 8;

__device__ __forceinline__ float relu(float v) { return v > 0.f ? v : 0.f; }

// ------------------------------------------------------------------ forward

__global__ void __launch_bounds__(256)
k_enc_points(const float4* __restrict__ X, int npts, const float* __restrict__ W,
             const __bf16* __restrict__ W2t, const __bf16* __restrict__ W3t,
             float* __restrict__ part_max, int* __restrict__ part_arg) {
  __shared__ __attribute__((aligned(16))) __bf16 H1[kRows * kH1S];
  __shared__ __attribute__((aligned(16))) __bf16 H2[kRows * kH2S];
  const int cloud = blockIdx.y, chunk = blockIdx.x, nchunks = gridDim.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int p0 = chunk * kRows;
  const float* W1 = W + kEncW1;
  const float* b1 = W + kEncB1;
  const float* b2 = W + kEncB2;
  const float* b3 = W + kEncB3;
  // layer 1 (FP32), kL1Threads threads per point, each a slice of the 64 outputs
  {
    constexpr int kL1Threads = 256 / kRows, kL1Cols = 64 / kL1Threads;
    const int row = t % kRows, j0 = (t / kRows) * kL1Cols;
    const int p = p0 + row;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool ok = p < npts;
    if (ok) x = X[(int64_t)cloud * npts + p];
#pragma unroll 8
    for (int jj = 0; jj < kL1Cols; ++jj) {
      const int j = j0 + jj;
      float a = b1[j];
      a = fmaf(x.x, W1[j], a);
      a = fmaf(x.y, W1[64 + j], a);
      a = fmaf(x.z, W1[128 + j], a);
      H1[row * kH1S + j] = (__bf16)(ok ? relu(a) : 0.f);
    }
  }
  __syncthreads();
  const int r16 = lane & 15, kq = (lane >> 4) * 8;  // fragment row/col and k offset
  // layer 2: wave w -> columns [32w, 32w + 32)
  {
    bf16x8 bf[2][2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
        bf[ct][ks] = *(const bf16x8*)(W2t + (32 * w + 16 * ct + r16) * 64 + 32 * ks + kq);
    float bias[2][1];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) bias[ct][0] = b2[32 * w + 16 * ct + r16];
    for (int rt = 0; rt < kRows / 16; ++rt) {
      f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const bf16x8 af = *(const bf16x8*)(H1 + (16 * rt + r16) * kH1S + 32 * ks + kq);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[ct][ks], acc[ct], 0, 0, 0);
      }
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rt + (lane >> 4) * 4 + e;
          const int col = 32 * w + 16 * ct + r16;
          H2[row * kH2S + col] = (__bf16)relu(acc[ct][e] + bias[ct][0]);
        }
    }
  }
  __syncthreads();
  // layer 3 + bias + ReLU + max-pool: wave w -> columns [64w, 64w + 64)
  {
    bf16x8 bf[4][4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        bf[ct][ks] = *(const bf16x8*)(W3t + (64 * w + 16 * ct + r16) * 128 + 32 * ks + kq);
    float bias[4], best[4];
    int brow[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      bias[ct] = b3[64 * w + 16 * ct + r16];
      best[ct] = -1.f;
      brow[ct] = 0x7fffffff;
    }
    const int valid = min(kRows, npts - p0);
    for (int rt = 0; rt < kRows / 16; ++rt) {
      f32x4 acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                      {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 af = *(const bf16x8*)(H2 + (16 * rt + r16) * kH2S + 32 * ks + kq);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[ct][ks], acc[ct], 0, 0, 0);
      }
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rt + (lane >> 4) * 4 + e;
          const float v = relu(acc[ct][e] + bias[ct]);
          if (row < valid && v > best[ct]) {
            best[ct] = v;
            brow[ct] = row;
          }
        }
    }
    // combine the 4 lane groups holding the same column (lanes l, l^16, l^32, l^48)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
#pragma unroll
      for (int off = 16; off <= 32; off <<= 1) {
        const float ov = __shfl_xor(best[ct], off, 64);
        const int orow = __shfl_xor(brow[ct], off, 64);
        if (ov > best[ct] || (ov == best[ct] && orow < brow[ct])) {
          best[ct] = ov;
          brow[ct] = orow;
        }
      }
      if (lane < 16) {
        const int col = 64 * w + 16 * ct + lane;
        const int64_t o = ((int64_t)cloud * nchunks + chunk) * 256 + col;
        part_max[o] = best[ct];
        part_arg[o] = brow[ct] == 0x7fffffff ? 0x7fffffff : p0 + brow[ct];
      }
    }
  }
}

// fold chunk partials: feature F[cloud][c] and its point ARG[cloud][c]
__global__ void k_enc_pool(const float* __restrict__ part_max, const int* __restrict__ part_arg,
                           int nchunks, float* __restrict__ F, int* __restrict__ ARG) {
  const int cloud = blockIdx.x, c = threadIdx.x;
  float best = -1.f;
  int arg = 0;
  for (int k = 0; k < nchunks; ++k) {
    const int64_t o = ((int64_t)cloud * nchunks + k) * 256 + c;
    const float v = part_max[o];
    if (v > best) {  // chunks in point order: the first maximum keeps the lowest index
      best = v;
      arg = part_arg[o];
    }
  }
  F[cloud * 256 + c] = best < 0.f ? 0.f : best;
  ARG[cloud * 256 + c] = arg == 0x7fffffff ? 0 : arg;
}

// head: one block (256 threads) per pair
__global__ void __launch_bounds__(256)
k_enc_head(const float* __restrict__ F, const float* __restrict__ W, const float* __restrict__ labels,
           int npairs, float* __restrict__ Z, float* __restrict__ OUT, float* __restrict__ DOUT,
           float* __restrict__ loss_part) {
  __shared__ float g[512];
  __shared__ float z[256];
  const int b = blockIdx.x, t = threadIdx.x;
  g[t] = F[(2 * b) * 256 + t];
  g[256 + t] = F[(2 * b + 1) * 256 + t];
  __syncthreads();
  const float* Hw1 = W + kEncHw1;
  float s = W[kEncHb1 + t];
  for (int i = 0; i < 512; ++i) s = fmaf(g[i], Hw1[i * 256 + t], s);
  z[t] = relu(s);
  Z[b * 256 + t] = z[t];
  __syncthreads();
  if (t < 6) {
    const float* Hw2 = W + kEncHw2;
    float o = W[kEncHb2 + t];
    for (int i = 0; i < 256; ++i) o = fmaf(z[i], Hw2[i * 6 + t], o);
    OUT[b * 6 + t] = o;
    if (labels) {
      const float d = o - labels[b * 6 + t];
      DOUT[b * 6 + t] = 2.f * d / (float)npairs;
      loss_part[b * 6 + t] = d * d / (float)npairs;
    }
  }
}

// dZ[b][j] = relu'(z) * sum_o dOut[b][o] Hw2[j][o];  dG[b][i] = sum_j dZ[b][j] Hw1[i][j]
__global__ void k_enc_head_back(const float* __restrict__ Z, const float* __restrict__ DOUT,
                                const float* __restrict__ W, float* __restrict__ DZ,
                                float* __restrict__ DG) {
  __shared__ float dz[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const float* Hw2 = W + kEncHw2;
  float s = 0.f;
#pragma unroll
  for (int o = 0; o < 6; ++o) s = fmaf(DOUT[b * 6 + o], Hw2[t * 6 + o], s);
  dz[t] = Z[b * 256 + t] > 0.f ? s : 0.f;
  DZ[b * 256 + t] = dz[t];
  __syncthreads();
  const float* Hw1 = W + kEncHw1;
  for (int i = t; i < 512; i += 256) {
    float a = 0.f;
    const float* row = Hw1 + i * 256;
    for (int j = 0; j < 256; ++j) a = fmaf(dz[j], row[j], a);
    DG[b * 512 + i] = a;
  }
}

// C[i][j] = sum_{r < R} A[r * lda + i] * B[r * ldb + j] (+ bias sums of B when
// i == M), one thread per output, rows in order: deterministic.  Used with R up to
// a few thousand (pairs); larger R goes through the two-stage version.
__global__ void k_outer_small(const float* __restrict__ A, int lda, const float* __restrict__ B,
                              int ldb, int R, int M, int N, float* __restrict__ C,
                              float* __restrict__ bias) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (M + 1) * N) return;
  const int i = idx / N, j = idx % N;
  float s = 0.f;
  if (i < M) {
    for (int r = 0; r < R; ++r) s = fmaf(A[(int64_t)r * lda + i], B[(int64_t)r * ldb + j], s);
    C[i * N + j] = s;
  } else {
    for (int r = 0; r < R; ++r) s += B[(int64_t)r * ldb + j];
    bias[j] = s;
  }
}

// stage 1 of a large outer-product sum: block k sums rows [k*rows, (k+1)*rows)
// into part[k][(M+1)*N] (last row = bias sums)
__global__ void k_outer_part(const float* __restrict__ A, int lda, const float* __restrict__ B,
                             int ldb, int R, int M, int N, int rows, float* __restrict__ part) {
  const int k = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (M + 1) * N) return;
  const int i = idx / N, j = idx % N;
  const int r0 = k * rows, r1 = min(R, r0 + rows);
  float s = 0.f;
  // unrolled so that the loads of several rows are in flight; one accumulator,
  // the rows in order
  if (i < M) {
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s = fmaf(A[(int64_t)r * lda + i], B[(int64_t)r * ldb + j], s);
  } else {
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += B[(int64_t)r * ldb + j];
  }
  part[(int64_t)k * (M + 1) * N + idx] = s;
}

// k_outer_part with a 4 x 4 block of outputs per thread (M, N, lda, ldb multiples
// of 4, A and B 16-byte aligned): per row one float4 of A and one of B feed 16
// multiply-adds, where one output per thread read two values per multiply-add and
// was bound by those loads.  Every output is the same fmaf chain over the rows in
// order, so the sums are bitwise those of k_outer_part.  Threads past the 4 x 4
// blocks sum the bias row, four columns each.
__global__ void k_outer_part4(const float* __restrict__ A, int lda, const float* __restrict__ B,
                              int ldb, int R, int M, int N, int rows, float* __restrict__ part) {
  const int k = blockIdx.y;
  const int nq = N >> 2, tiles = (M >> 2) * nq;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles + nq) return;
  const int r0 = k * rows, r1 = min(R, r0 + rows);
  float* out = part + (int64_t)k * (M + 1) * N;
  if (t < tiles) {
    const int i0 = (t / nq) * 4, j0 = (t % nq) * 4;
    float acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = 0.f;
#pragma unroll 4
    for (int r = r0; r < r1; ++r) {
      const float4 a = *reinterpret_cast<const float4*>(A + (int64_t)r * lda + i0);
      const float4 b = *reinterpret_cast<const float4*>(B + (int64_t)r * ldb + j0);
      const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fmaf(av[u], bv[v], acc[u][v]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) out[(i0 + u) * N + j0 + v] = acc[u][v];
  } else {
    const int j0 = (t - tiles) * 4;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int r = r0; r < r1; ++r) {
      const float4 b = *reinterpret_cast<const float4*>(B + (int64_t)r * ldb + j0);
      acc[0] += b.x;
      acc[1] += b.y;
      acc[2] += b.z;
      acc[3] += b.w;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) out[M * N + j0 + v] = acc[v];
  }
}

__global__ void k_outer_fold(const float* __restrict__ part, int nparts, int M, int N,
                             float* __restrict__ C, float* __restrict__ bias) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (M + 1) * N) return;
  float s = 0.f;
  for (int k = 0; k < nparts; ++k) s += part[(int64_t)k * (M + 1) * N + idx];
  if (idx < M * N)
    C[idx] = s;
  else
    bias[idx - M * N] = s;
}

// Per (cloud, channel) row of the max-pool backward: recompute h1, h2 at the
// channel's argmax point and push the channel's gradient down to layer 1.
// One block = 64 channels of one cloud.  Writes per row: x (3), h1 (64), da1 (64),
// h2 (128), da2 (128) and the row's df (used by the W3 gradient).
__global__ void __launch_bounds__(256)
k_enc_rows(const float4* __restrict__ X, int npts, const float* __restrict__ F,
           const int* __restrict__ ARG, const float* __restrict__ DG, const float* __restrict__ W,
           const __bf16* __restrict__ W2t, const __bf16* __restrict__ W3t, float* __restrict__ XA,
           float* __restrict__ H1A, float* __restrict__ DA1, float* __restrict__ H2A,
           float* __restrict__ DA2, float* __restrict__ DF) {
  __shared__ float h1[64][65];
  __shared__ float m1[64][64];
  __shared__ float da2s[64][129];
  __shared__ float xs[64][3];
  __shared__ float dfs[64];
  const int cloud = blockIdx.y, c0 = blockIdx.x * 64, t = threadIdx.x;
  const int pair = cloud >> 1, side = cloud & 1;
  const int64_t rbase = (int64_t)cloud * 256 + c0;
  if (t < 64) {
    const int c = c0 + t;
    const float f = F[cloud * 256 + c];
    const float df = f > 0.f ? DG[pair * 512 + side * 256 + c] : 0.f;
    dfs[t] = df;
    const float4 x = X[(int64_t)cloud * npts + ARG[cloud * 256 + c]];
    xs[t][0] = x.x;
    xs[t][1] = x.y;
    xs[t][2] = x.z;
    DF[rbase + t] = df;
    XA[(rbase + t) * 3 + 0] = x.x;
    XA[(rbase + t) * 3 + 1] = x.y;
    XA[(rbase + t) * 3 + 2] = x.z;
  }
  __syncthreads();
  // layer 1 for the 64 rows: thread -> (row r = t / 4, 16 columns)
  {
    const int r = t >> 2, j0 = (t & 3) * 16;
    for (int j = j0; j < j0 + 16; ++j) {
      float a = W[kEncB1 + j];
      a = fmaf(xs[r][0], W[kEncW1 + j], a);
      a = fmaf(xs[r][1], W[kEncW1 + 64 + j], a);
      a = fmaf(xs[r][2], W[kEncW1 + 128 + j], a);
      h1[r][j] = a > 0.f ? (float)(__bf16)a : 0.f;
      m1[r][j] = a > 0.f ? 1.f : 0.f;
      H1A[(rbase + r) * 64 + j] = h1[r][j];
    }
  }
  __syncthreads();
  // layer 2: thread -> column j = t % 128, rows r = t / 128 + 2m
  {
    const int j = t & 127;
    float w2[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) w2[i] = (float)W2t[j * 64 + i];
    const float b2 = W[kEncB2 + j];
    for (int r = t >> 7; r < 64; r += 2) {
      float a = b2;
#pragma unroll
      for (int i = 0; i < 64; ++i) a = fmaf(h1[r][i], w2[i], a);
      const float h2 = a > 0.f ? (float)(__bf16)a : 0.f;
      const float g = a > 0.f ? dfs[r] * (float)W3t[(c0 + r) * 128 + j] : 0.f;
      da2s[r][j] = g;
      H2A[(rbase + r) * 128 + j] = h2;
      DA2[(rbase + r) * 128 + j] = g;
    }
  }
  __syncthreads();
  // dh1 = da2 W2^T, masked: thread -> (row r = t / 4, 16 columns)
  {
    const int r = t >> 2, i0 = (t & 3) * 16;
    for (int i = i0; i < i0 + 16; ++i) {
      float s = 0.f;
      for (int j = 0; j < 128; ++j) s = fmaf(da2s[r][j], (float)W2t[j * 64 + i], s);
      DA1[(rbase + r) * 64 + i] = s * m1[r][i];
    }
  }
}

// dW3[k][c] = sum_clouds H2A[(cloud, c)][k] * DF[(cloud, c)]; db3[c] = sum DF
__global__ void k_enc_dw3(const float* __restrict__ H2A, const float* __restrict__ DF, int nclouds,
                          float* __restrict__ G) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 129 * 256) return;
  const int k = idx / 256, c = idx % 256;
  float s = 0.f;
  if (k < 128) {
    for (int cl = 0; cl < nclouds; ++cl) {
      const int64_t r = (int64_t)cl * 256 + c;
      s = fmaf(H2A[r * 128 + k], DF[r], s);
    }
    G[kEncW3 + k * 256 + c] = s;
  } else {
    for (int cl = 0; cl < nclouds; ++cl) s += DF[(int64_t)cl * 256 + c];
    G[kEncB3 + c] = s;
  }
}

// Adam (FP32 master weights), then refresh the BF16 transposed copies of W2, W3
__global__ void k_adam(float* __restrict__ W, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  W[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_enc_bf16_weights(const float* __restrict__ W, __bf16* __restrict__ W2t,
                                   __bf16* __restrict__ W3t) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < 64 * 128) {
    const int i = idx / 128, j = idx % 128;  // W2[i][j] -> W2t[j][i]
    W2t[j * 64 + i] = (__bf16)W[kEncW2 + idx];
  }
  if (idx < 128 * 256) {
    const int k = idx / 256, c = idx % 256;  // W3[k][c] -> W3t[c][k]
    W3t[c * 128 + k] = (__bf16)W[kEncW3 + idx];
  }
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int enc_forward_chunks(int npts) { return (npts + kRows - 1) / kRows; }

void launch_enc_bf16_weights(const float* W, __bf16* W2t, __bf16* W3t, hipStream_t s) {
  hipLaunchKernelGGL(k_enc_bf16_weights, dim3(nb(128 * 256, 256)), dim3(256), 0, s, W, W2t, W3t);
}

void launch_enc_forward(const EncBuffers& e, int npairs, int npts, const float* labels,
                        hipStream_t s) {
  const int nclouds = 2 * npairs;
  const int nchunks = enc_forward_chunks(npts);
  hipLaunchKernelGGL(k_enc_points, dim3(nchunks, nclouds), dim3(256), 0, s, e.X, npts, e.W, e.W2t,
                     e.W3t, e.part_max, e.part_arg);
  hipLaunchKernelGGL(k_enc_pool, dim3(nclouds), dim3(256), 0, s, e.part_max, e.part_arg, nchunks,
                     e.F, e.ARG);
  hipLaunchKernelGGL(k_enc_head, dim3(npairs), dim3(256), 0, s, e.F, e.W, labels, npairs, e.Z,
                     e.OUT, e.DOUT, e.loss_part);
}

void launch_enc_backward(const EncBuffers& e, int npairs, int npts, hipStream_t s) {
  const int nclouds = 2 * npairs;
  float* G = e.G;
  hipLaunchKernelGGL(k_enc_head_back, dim3(npairs), dim3(256), 0, s, e.Z, e.DOUT, e.W, e.DZ, e.DG);
  // head weights: dHw2 = Z^T dOut, dHw1 = G^T dZ (G = pooled features)
  hipLaunchKernelGGL(k_outer_small, dim3(nb(257 * 6, 256)), dim3(256), 0, s, e.Z, 256, e.DOUT, 6,
                     npairs, 256, 6, G + kEncHw2, G + kEncHb2);
  hipLaunchKernelGGL(k_outer_small, dim3(nb(513 * 256, 256)), dim3(256), 0, s, e.F, 512, e.DZ, 256,
                     npairs, 512, 256, G + kEncHw1, G + kEncHb1);
  // per (cloud, channel) rows
  hipLaunchKernelGGL(k_enc_rows, dim3(4, nclouds), dim3(256), 0, s, e.X, npts, e.F, e.ARG, e.DG,
                     e.W, e.W2t, e.W3t, e.XA, e.H1A, e.DA1, e.H2A, e.DA2, e.DF);
  hipLaunchKernelGGL(k_enc_dw3, dim3(nb(129 * 256, 256)), dim3(256), 0, s, e.H2A, e.DF, nclouds, G);
  const int R = nclouds * 256;
  const int rows = 1024;
  const int parts = (R + rows - 1) / rows;
  hipLaunchKernelGGL(k_outer_part4, dim3(nb(16 * 32 + 32, 256), parts), dim3(256), 0, s, e.H1A, 64,
                     e.DA2, 128, R, 64, 128, rows, e.part_grad);
  hipLaunchKernelGGL(k_outer_fold, dim3(nb(65 * 128, 256)), dim3(256), 0, s, e.part_grad, parts, 64,
                     128, G + kEncW2, G + kEncB2);
  hipLaunchKernelGGL(k_outer_part, dim3(nb(4 * 64, 256), parts), dim3(256), 0, s, e.XA, 3, e.DA1,
                     64, R, 3, 64, rows, e.part_grad);
  hipLaunchKernelGGL(k_outer_fold, dim3(nb(4 * 64, 256)), dim3(256), 0, s, e.part_grad, parts, 3, 64,
                     G + kEncW1, G + kEncB1);
}

void launch_adam(float* W, const float* G, float* M, float* V, int n, float lr, int step,
                 hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nb(n, 256)), dim3(256), 0, s, W, G, M, V, n, lr, b1, b2, eps, c1,
                     c2);
}

int64_t enc_part_grad_floats(int npairs) {
  const int R = 2 * npairs * 256;
  const int parts = (R + 1023) / 1024;
  return (int64_t)parts * 65 * 128;
}

}  // namespace lcr
