// Learned warm start for the PCG (extension; no reference counterpart): a message-passing
// network over the tet face-adjacency graph that predicts dlambda of the first Newton
// iteration from per-constraint features. Forward and backward run here; the three
// 128 -> 64 message-passing layers use BF16 MFMA (gfx950 v_mfma_f32_32x32x16_bf16) with FP32
// accumulation, the 8 -> 64 encoder and 64 -> 1 decoder are FP32.
//
// Network (n tets, H = 64):
//   h0 = relu(x W_in + b_in)                         x: n x 8 features
//   m_l = mean_{j in N(i)} h_{l-1,j}                 N = face neighbours (<= 4)
//   h_l = relu(h_{l-1} W_l^self + m_l W_l^nbr + b_l)   l = 1..3
//   y   = h_3 w_out + b_out
// Parameters (FP32, flattened): W_in[8][64], b_in[64], {W_l[128][64] (rows 0..63 self,
// 64..127 neighbour), b_l[64]} x 3, w_out[64], b_out[1].
// Activations are stored as BF16 (n x 64 row-major); BF16 weight copies are kept both
// transposed (out-major, the forward B operand) and as-is (the backward B operand).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "mgpbd_internal.h"

namespace mgpbd_gnn {

constexpr int H = MGPBD_GNN_HIDDEN;       // 64
constexpr int F = MGPBD_GNN_FEATURES;     // 8
constexpr int NL = MGPBD_GNN_LAYERS;      // 3
constexpr int kBlock = 256;

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v16f32 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// offsets into the flattened FP32 parameter vector
__host__ __device__ constexpr int off_win() { return 0; }
__host__ __device__ constexpr int off_bin() { return F * H; }
__host__ __device__ constexpr int off_wl(int l) { return F * H + H + l * (2 * H * H + H); }
__host__ __device__ constexpr int off_bl(int l) { return off_wl(l) + 2 * H * H; }
__host__ __device__ constexpr int off_wout() { return off_wl(NL); }
__host__ __device__ constexpr int off_bout() { return off_wout() + H; }
__host__ __device__ constexpr int n_params() { return off_bout() + 1; }

// ------------------------------------------------------------------------------------------
// MFMA GEMM: Y[n x 64] = act(A1 W1 + A2 W2 + bias), A* BF16 n x 64, W* as out-major BF16
// [64 out][64 in]. One wavefront computes 32 rows x 64 columns (two 32x32 MFMA tiles).
// gfx950's v_mfma_f32_32x32x16_bf16 operand layout (wave64): A fragment lane l holds
// A[l%32][8(l/32) .. +7], B fragment holds B[8(l/32) .. +7][l%32]; result register r of lane
// l is D[8(r/4) + 4(l/32) + r%4][l%32]. A 64-long inner dimension is four instructions.
// ------------------------------------------------------------------------------------------
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef short v8i16 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v16f32 mfma16(v8i16 a, v8i16 b, v16f32 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf16, a),
                                                 __builtin_bit_cast(v8bf16, b), c, 0, 0, 0);
}
__device__ __forceinline__ v4i16 load4(const uint16_t* p) {
  return *reinterpret_cast<const v4i16*>(p);
}
// eight BF16 values at a 16-byte aligned address (one 16-byte load)
__device__ __forceinline__ v8i16 load8(const uint16_t* p) {
  return *reinterpret_cast<const v8i16*>(p);
}
// eight BF16 values at an 8-byte aligned address (LDS rows padded to 68 halves): two loads
__device__ __forceinline__ v8i16 load8_lds(const uint16_t* p) {
  const v4i16 lo = load4(p), hi = load4(p + 4);
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__global__ void __launch_bounds__(kBlock)
k_gemm64(int n, const uint16_t* __restrict__ A1, const uint16_t* __restrict__ W1t,
         const uint16_t* __restrict__ A2, const uint16_t* __restrict__ W2t,
         const float* __restrict__ bias, int relu, uint16_t* __restrict__ Ybf,
         float* __restrict__ Yf) {
  const int lane = threadIdx.x & 63;
  const int base = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 32;
  if (base >= n) return;
  const int ra = base + (lane & 31);
  const int kq = 8 * (lane >> 5);
  v16f32 acc0 = {}, acc1 = {};
  const v8i16 zero = {};
#pragma unroll
  for (int s = 0; s < H / 16; ++s) {
    const v8i16 a = ra < n ? load8(A1 + (size_t)ra * H + 16 * s + kq) : zero;
    const v8i16 b0 = load8(W1t + (size_t)(lane & 31) * H + 16 * s + kq);
    const v8i16 b1 = load8(W1t + (size_t)(32 + (lane & 31)) * H + 16 * s + kq);
    acc0 = mfma16(a, b0, acc0);
    acc1 = mfma16(a, b1, acc1);
  }
  if (A2) {
#pragma unroll
    for (int s = 0; s < H / 16; ++s) {
      const v8i16 a = ra < n ? load8(A2 + (size_t)ra * H + 16 * s + kq) : zero;
      const v8i16 b0 = load8(W2t + (size_t)(lane & 31) * H + 16 * s + kq);
      const v8i16 b1 = load8(W2t + (size_t)(32 + (lane & 31)) * H + 16 * s + kq);
      acc0 = mfma16(a, b0, acc0);
      acc1 = mfma16(a, b1, acc1);
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = base + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
    if (row >= n) continue;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = 32 * c + (lane & 31);
      float v = (c == 0 ? acc0[r] : acc1[r]) + (bias ? bias[col] : 0.0f);
      if (relu) v = fmaxf(v, 0.0f);
      if (Ybf) Ybf[(size_t)row * H + col] = f2bf(v);
      if (Yf) Yf[(size_t)row * H + col] = v;
    }
  }
}

// Both input-gradient products of a layer from one read of dZ: Y1 (BF16) = A W1 and
// Y2 (FP32) = A W2, each with the MFMA sequence of k_gemm64 (bit-identical to two calls).
__global__ void __launch_bounds__(kBlock)
k_gemm64_dual(int n, const uint16_t* __restrict__ A, const uint16_t* __restrict__ W1t,
              const uint16_t* __restrict__ W2t, uint16_t* __restrict__ Y1bf,
              float* __restrict__ Y2f) {
  const int lane = threadIdx.x & 63;
  const int base = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 32;
  if (base >= n) return;
  const int ra = base + (lane & 31);
  const int kq = 8 * (lane >> 5);
  v16f32 acc0 = {}, acc1 = {}, acc2 = {}, acc3 = {};
  const v8i16 zero = {};
#pragma unroll
  for (int s = 0; s < H / 16; ++s) {
    const v8i16 a = ra < n ? load8(A + (size_t)ra * H + 16 * s + kq) : zero;
    const v8i16 b0 = load8(W1t + (size_t)(lane & 31) * H + 16 * s + kq);
    const v8i16 b1 = load8(W1t + (size_t)(32 + (lane & 31)) * H + 16 * s + kq);
    const v8i16 b2 = load8(W2t + (size_t)(lane & 31) * H + 16 * s + kq);
    const v8i16 b3 = load8(W2t + (size_t)(32 + (lane & 31)) * H + 16 * s + kq);
    acc0 = mfma16(a, b0, acc0);
    acc1 = mfma16(a, b1, acc1);
    acc2 = mfma16(a, b2, acc2);
    acc3 = mfma16(a, b3, acc3);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = base + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
    if (row >= n) continue;
    const int col = lane & 31;
    Y1bf[(size_t)row * H + col] = f2bf(acc0[r] + 0.0f);
    Y1bf[(size_t)row * H + 32 + col] = f2bf(acc1[r] + 0.0f);
    Y2f[(size_t)row * H + col] = acc2[r] + 0.0f;
    Y2f[(size_t)row * H + 32 + col] = acc3[r] + 0.0f;
  }
}

// Inference layer with the neighbour mean fused in: Y = relu(H W_self + mean_nbr(H) W_nbr + b)
// over 32-row MFMA tiles, with the arithmetic of a mean pass followed by k_gemm64 (same fragments, same
// MFMA sequence), so the output equals theirs bit for bit without the mean's n x 64 round trip
// through memory. Memory side: each block stages both weight matrices in LDS once and its
// wavefronts loop over tiles; a tile's own rows are loaded as 16-byte row chunks (eight lanes
// per 128-byte row), and each (row, 8-column chunk) pair gathers the same chunk of its four
// neighbour rows and forms the mean there (the four rows summed in FP32 in neighbour order,
// times 1/deg, rounded to BF16). MFMA operands are read from LDS rows padded to 68 halves
// (8-byte aligned, rows two banks apart), and the output tile goes back through LDS to
// 16-byte row-chunk stores.
constexpr int kLdsRow = H + 4;
constexpr int kTileWaves = kBlock / 64;
constexpr int kLayerBlocks = 768;   // 3 blocks (52 KiB of LDS each) per CU on 256 CUs
constexpr int kEncoderBlocks = 2048; // encoder grid: 16 rows per lane group at 1M rows

typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void lds_store8(uint16_t* dst, v8u16 v) {
  reinterpret_cast<v4i16*>(dst)[0] = v4i16{(short)v[0], (short)v[1], (short)v[2], (short)v[3]};
  reinterpret_cast<v4i16*>(dst)[1] = v4i16{(short)v[4], (short)v[5], (short)v[6], (short)v[7]};
}
__device__ __forceinline__ v8u16 lds_load8(const uint16_t* src) {
  const v4i16 lo = reinterpret_cast<const v4i16*>(src)[0];
  const v4i16 hi = reinterpret_cast<const v4i16*>(src)[1];
  return v8u16{(unsigned short)lo.x, (unsigned short)lo.y, (unsigned short)lo.z, (unsigned short)lo.w,
               (unsigned short)hi.x, (unsigned short)hi.y, (unsigned short)hi.z, (unsigned short)hi.w};
}

// Decoder fused into the last inference layer: with dlam set, the layer's output rows (BF16, in
// LDS) go straight through k_decoder_predict's arithmetic instead of to memory.
struct FusedDecoder {
  const float* P;     // parameters (w_out, b_out)
  const float* lp;    // previous lambda (nullable)
  const float* st;    // feature statistics (st[0]: the prediction scale)
  float* y;           // decoder output (nullable)
  float* dlam;        // prediction; nullptr: no fused decoder
};

__global__ void __launch_bounds__(kBlock)
k_layer_mean(int n, const int32_t* __restrict__ nbr, const uint16_t* __restrict__ Hin,
             const uint16_t* __restrict__ Wst, const uint16_t* __restrict__ Wnt,
             const float* __restrict__ bias, uint16_t* __restrict__ Hout,
             uint16_t* __restrict__ Mout, FusedDecoder dec) {
  __shared__ __align__(16) uint16_t wl[2][H * kLdsRow];
  __shared__ __align__(16) uint16_t tl[kTileWaves][2][32 * kLdsRow];
  __shared__ float wo[H];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = threadIdx.x; c < 2 * H * 8; c += kBlock) {   // 16-byte chunks of both matrices
    const int m = c / (H * 8), r = (c / 8) % H, ch = c % 8;
    const uint16_t* src = (m ? Wnt : Wst) + r * H + 8 * ch;
    lds_store8(wl[m] + r * kLdsRow + 8 * ch, *reinterpret_cast<const v8u16*>(src));
  }
  if (dec.dlam && threadIdx.x < H) wo[threadIdx.x] = dec.P[off_wout() + threadIdx.x];
  const int n_tiles = (n + 31) / 32;
  const int kq = 8 * (lane >> 5), rl = lane & 31;
  uint16_t* ts = tl[w][0];
  uint16_t* tm = tl[w][1];
  const float bias0 = bias[rl], bias1 = bias[32 + rl];
  // the blocks sweep the tiles as one front (tile groups dealt round-robin), so the face
  // neighbours of the rows in flight (a few to a few hundred rows away, ~20k for the third
  // axis of the block meshes) were mostly just read by another block; the loop bound is the
  // block's, so every wavefront meets every barrier
  for (int t0 = blockIdx.x * kTileWaves; t0 < n_tiles; t0 += gridDim.x * kTileWaves) {
    const int base = (t0 + w) * 32;
    // all loads of the tile are issued before any is used: rows past n read row n - 1 and
    // absent neighbours re-read the row itself, and selects drop them (a dropped neighbour
    // leaves the sum untouched, as skipping it does)
    int rows[4];
    int4 nq[4];
    v8u16 hv[4], hn[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lane + 64 * q;
      rows[q] = min(base + (c >> 3), n - 1);
      hv[q] = *reinterpret_cast<const v8u16*>(Hin + (size_t)rows[q] * H + 8 * (c & 7));
      nq[q] = reinterpret_cast<const int4*>(nbr)[rows[q]];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ch = (lane + 64 * q) & 7;
      const int nb4[4] = {nq[q].x, nq[q].y, nq[q].z, nq[q].w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = nb4[k] >= 0 ? nb4[k] : rows[q];
        hn[q][k] = *reinterpret_cast<const v8u16*>(Hin + (size_t)j * H + 8 * ch);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lane + 64 * q, r = c >> 3, ch = c & 7;
      const bool live = base + r < n;
      const int nb4[4] = {nq[q].x, nq[q].y, nq[q].z, nq[q].w};
      int d = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) d += nb4[k] >= 0;
      const float inv = d > 0 ? 1.0f / (float)d : 0.0f;
      float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float t = m[e] + bf2f(hn[q][k][e]);
          m[e] = nb4[k] >= 0 ? t : m[e];
        }
      }
      v8u16 mv = {}, hz = {};
#pragma unroll
      for (int e = 0; e < 8; ++e) mv[e] = f2bf(m[e] * inv);
      if (Mout && live)   // the training forward keeps the mean for the weight gradient
        *reinterpret_cast<v8u16*>(Mout + (size_t)(base + r) * H + 8 * ch) = mv;
      lds_store8(ts + r * kLdsRow + 8 * ch, live ? hv[q] : hz);
      lds_store8(tm + r * kLdsRow + 8 * ch, live ? mv : hz);
    }
    __syncthreads();
    v16f32 acc0 = {}, acc1 = {};
#pragma unroll
    for (int s = 0; s < H / 16; ++s) {
      const v8i16 a = load8_lds(ts + rl * kLdsRow + 16 * s + kq);
      const v8i16 b0 = load8_lds(wl[0] + rl * kLdsRow + 16 * s + kq);
      const v8i16 b1 = load8_lds(wl[0] + (32 + rl) * kLdsRow + 16 * s + kq);
      acc0 = mfma16(a, b0, acc0);
      acc1 = mfma16(a, b1, acc1);
    }
#pragma unroll
    for (int s = 0; s < H / 16; ++s) {
      const v8i16 a = load8_lds(tm + rl * kLdsRow + 16 * s + kq);
      const v8i16 b0 = load8_lds(wl[1] + rl * kLdsRow + 16 * s + kq);
      const v8i16 b1 = load8_lds(wl[1] + (32 + rl) * kLdsRow + 16 * s + kq);
      acc0 = mfma16(a, b0, acc0);
      acc1 = mfma16(a, b1, acc1);
    }
    __syncthreads();   // every fragment read of the tile is done before it is overwritten
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rr = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
      ts[rr * kLdsRow + rl] = f2bf(fmaxf(acc0[r] + bias0, 0.0f));
      ts[rr * kLdsRow + 32 + rl] = f2bf(fmaxf(acc1[r] + bias1, 0.0f));
    }
    __syncthreads();
    if (dec.dlam) {
      // one lane per row: k_decoder_predict's products and pairwise sums over the row's
      // BF16 values (c + 32, then c + 16, ...)
      const int row = base + rl;
      if (lane < 32 && row < n) {
        float p[H];
#pragma unroll
        for (int c = 0; c < H; ++c) p[c] = bf2f(ts[rl * kLdsRow + c]) * wo[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
          for (int c = 0; c < off; ++c) p[c] = p[c] + p[c + off];
        const float yi = p[0] + dec.P[off_bout()];
        if (dec.y) dec.y[row] = yi;
        dec.dlam[row] = (dec.lp ? dec.lp[row] : 0.0f) + dec.st[0] * yi;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = lane + 64 * q, r = c >> 3, ch = c & 7, row = base + r;
        if (row < n)
          *reinterpret_cast<v8u16*>(Hout + (size_t)row * H + 8 * ch) = lds_load8(ts + r * kLdsRow + 8 * ch);
      }
    }
    __syncthreads();   // the tile's LDS rows are free for the next tile
  }
}

// Weight gradient partials: P[block][k][j] = sum over the block's rows i of A[i][k] dZ[i][j]
// (A, dZ BF16 n x 64). Each of the 4 wavefronts owns one 32x32 quadrant of the 64x64 result;
// the MFMA inner dimension runs over rows.
constexpr int kWgradRows = 1024;
constexpr int kColRows = 256;   // rows per block of the column-sum and encoder-gradient passes

// With A2 (nullable) the same block also forms P2 = A2^T dZ from the dZ fragments it loaded
// for P, and with Pb (nullable) the column sums of dZ over its rows (the bias gradient): the
// two waves with qk = 0 add the dZ values of their fragments (each lane eight rows of one
// column), and lanes l and l + 32 (the two row halves) combine in a fixed order. The rows go
// through LDS 64 at a time (16-byte row-chunk loads, rows padded to 68 halves); the MFMA
// sequence over 16-row steps is the same as reading the fragments from memory.
constexpr int kWgChunk = 64;
__global__ void __launch_bounds__(kBlock)
k_wgrad(int n, const uint16_t* __restrict__ A, const uint16_t* __restrict__ dZ,
        float* __restrict__ partial, const uint16_t* __restrict__ A2,
        float* __restrict__ partial2, float* __restrict__ Pb) {
  __shared__ __align__(16) uint16_t sa[kWgChunk * kLdsRow];
  __shared__ __align__(16) uint16_t sa2[kWgChunk * kLdsRow];
  __shared__ __align__(16) uint16_t sz[kWgChunk * kLdsRow];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int qk = w >> 1, qj = w & 1;
  const int r0 = blockIdx.x * kWgradRows;
  const int r1 = min(n, r0 + kWgradRows);
  const int kq = 8 * (lane >> 5);
  const int ca = 32 * qk + (lane & 31), cz = 32 * qj + (lane & 31);
  v16f32 acc = {}, acc2 = {};
  float cs = 0.0f;
  const v8u16 zero8 = {};
  for (int c0 = r0; c0 < r1; c0 += kWgChunk) {
    for (int q = threadIdx.x; q < kWgChunk * 8; q += kBlock) {
      const int rr = q >> 3, ch = q & 7, row = c0 + rr;
      const bool in = row < r1;
      lds_store8(sa + rr * kLdsRow + 8 * ch,
                 in ? *reinterpret_cast<const v8u16*>(A + (size_t)row * H + 8 * ch) : zero8);
      lds_store8(sz + rr * kLdsRow + 8 * ch,
                 in ? *reinterpret_cast<const v8u16*>(dZ + (size_t)row * H + 8 * ch) : zero8);
      if (A2)
        lds_store8(sa2 + rr * kLdsRow + 8 * ch,
                   in ? *reinterpret_cast<const v8u16*>(A2 + (size_t)row * H + 8 * ch) : zero8);
    }
    __syncthreads();
    const int rows = min(kWgChunk, r1 - c0);
    // rows past r1 in the chunk are zero in LDS, so a 16-row step may run past `rows`
    for (int s0 = 0; s0 < rows; s0 += 16) {
      const int i = s0 + kq;
      v8i16 a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[e] = (short)sa[(i + e) * kLdsRow + ca];
        b[e] = (short)sz[(i + e) * kLdsRow + cz];
      }
      acc = mfma16(a, b, acc);
      if (A2) {
        v8i16 a2;
#pragma unroll
        for (int e = 0; e < 8; ++e) a2[e] = (short)sa2[(i + e) * kLdsRow + ca];
        acc2 = mfma16(a2, b, acc2);
      }
      if (Pb && qk == 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) cs += bf2f((uint16_t)b[e]);
      }
    }
    __syncthreads();   // the chunk's LDS rows are free for the next chunk
  }
  float* P = partial + (size_t)blockIdx.x * H * H;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int k = 32 * qk + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
    const int j = 32 * qj + (lane & 31);
    P[k * H + j] = acc[r];
  }
  if (A2) {
    float* P2 = partial2 + (size_t)blockIdx.x * H * H;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int k = 32 * qk + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
      const int j = 32 * qj + (lane & 31);
      P2[k * H + j] = acc2[r];
    }
  }
  if (Pb && qk == 0) {
    const float other = __shfl_xor(cs, 32, 64);
    if (lane < 32) Pb[(size_t)blockIdx.x * H + 32 * qj + lane] = cs + other;
  }
}

__global__ void __launch_bounds__(kBlock)
k_sum_partials(int nblk, int len, int slice, int stride, const float* partial, float* out) {
  // block (x, y): columns [64x, 64x + 64) summed over rows r in [y slice, min(nblk, (y + 1)
  // slice)), row r at partial + r stride len; 256 / 64 = 4 row groups (fixed order), then the
  // groups in order. The result goes to out + y slice stride len (in place into the slice's
  // first row when out == partial).
  __shared__ float sh[kBlock];
  const int cols = min(len, 64);
  const int G = kBlock / cols;
  const int c = threadIdx.x % cols, g = threadIdx.x / cols;
  const int e = blockIdx.x * 64 + c;
  const int r0 = blockIdx.y * slice, r1 = min(nblk, r0 + slice);
  float s = 0.0f;
  if (g < G && e < len)
    for (int r = r0 + g; r < r1; r += G) s += partial[(size_t)r * stride * len + e];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cols && e < len) {
    float t = 0.0f;
    for (int q = 0; q < G; ++q) t += sh[q * cols + threadIdx.x];
    out[(size_t)r0 * stride * len + e] = t;
  }
}

// out[0 .. len) = sum over the nblk rows of partial (nblk x len, row-major), deterministic:
// slices of 32 rows summed in place into each slice's first row by parallel blocks, then those
// rows summed (a single pass when nblk <= 32). partial is overwritten.
static void sum_partials(hipStream_t s, int nblk, int len, float* partial, float* out) {
  const int gx = (len + 63) / 64;
  constexpr int kSlice = 32;
  if (nblk <= kSlice) {
    k_sum_partials<<<dim3(gx, 1), kBlock, 0, s>>>(nblk, len, nblk, 1, partial, out);
    return;
  }
  const int ns = (nblk + kSlice - 1) / kSlice;
  k_sum_partials<<<dim3(gx, ns), kBlock, 0, s>>>(nblk, len, kSlice, 1, partial, partial);
  k_sum_partials<<<dim3(gx, 1), kBlock, 0, s>>>(ns, len, ns, kSlice, partial, out);
}

// ------------------------------------------------------------------------------------------
// Elementwise / small kernels
// ------------------------------------------------------------------------------------------
// Backward of the mean and of the ReLU below it: with dh_j = dh_j + sum_{i in N(j)} dm_i /
// deg_i (the face graph is symmetric; FP32), dz_j = dh_j where h_j > 0, else 0, stored BF16
// (the next layer down's dZ). Eight columns per thread (16-byte BF16 loads and stores).
__global__ void __launch_bounds__(kBlock)
k_mean_nbr_relu_bwd(int n, const int32_t* __restrict__ nbr, const uint16_t* __restrict__ dm,
                    const float* __restrict__ dh, const uint16_t* __restrict__ h,
                    uint16_t* __restrict__ dz) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int j = t >> 3;
  const int c = 8 * (t & 7);
  if (j >= n) return;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int4 nj = *reinterpret_cast<const int4*>(nbr + 4 * (size_t)j);
  const int ids[4] = {nj.x, nj.y, nj.z, nj.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = ids[q];
    if (i < 0) continue;
    const int4 ni = *reinterpret_cast<const int4*>(nbr + 4 * (size_t)i);
    const int d = (ni.x >= 0) + (ni.y >= 0) + (ni.z >= 0) + (ni.w >= 0);
    const float inv = 1.0f / (float)d;
    const uint4 v = *reinterpret_cast<const uint4*>(dm + (size_t)i * H + c);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      s[2 * e] = fmaf(bf2f((uint16_t)w[e]), inv, s[2 * e]);
      s[2 * e + 1] = fmaf(bf2f((uint16_t)(w[e] >> 16)), inv, s[2 * e + 1]);
    }
  }
  const float4 a = *reinterpret_cast<const float4*>(dh + (size_t)j * H + c);
  const float4 b = *reinterpret_cast<const float4*>(dh + (size_t)j * H + c + 4);
  const float g[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  const uint4 hv = *reinterpret_cast<const uint4*>(h + (size_t)j * H + c);
  const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
  uint32_t o[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float lo = bf2f((uint16_t)hw[e]) > 0.0f ? g[2 * e] + s[2 * e] : 0.0f;
    const float hi = bf2f((uint16_t)(hw[e] >> 16)) > 0.0f ? g[2 * e + 1] + s[2 * e + 1] : 0.0f;
    o[e] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
  }
  *reinterpret_cast<uint4*>(dz + (size_t)j * H + c) = make_uint4(o[0], o[1], o[2], o[3]);
}

// h0 = relu(x W_in + b_in) (FP32 weights, fma over the 8 features in order), 8 channels per
// thread, one 16-byte BF16 store
__global__ void __launch_bounds__(kBlock)
k_encoder8(int n, const float* __restrict__ x, const float* __restrict__ P,
           uint16_t* __restrict__ h0) {
  // eight lanes per row, lane q owning output columns 8q .. 8q + 7: its 8 x 8 weights and 8
  // biases stay in registers for all of its rows (grid-stride over rows), so a row costs two
  // 16-byte feature loads, 64 FMAs and one 16-byte store per lane
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int c0 = 8 * (t & 7);
  float w[F][8], bias[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    bias[e] = P[off_bin() + c0 + e];
#pragma unroll
    for (int k = 0; k < F; ++k) w[k][e] = P[off_win() + k * H + c0 + e];
  }
  const int stride = (gridDim.x * kBlock) >> 3;
  for (int i = t >> 3; i < n; i += stride) {
    const float4 xa = reinterpret_cast<const float4*>(x)[2 * (size_t)i];
    const float4 xb = reinterpret_cast<const float4*>(x)[2 * (size_t)i + 1];
    const float xi[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
    uint16_t out[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float v = bias[e];
#pragma unroll
      for (int k = 0; k < F; ++k) v = fmaf(xi[k], w[k][e], v);
      out[e] = f2bf(fmaxf(v, 0.0f));
    }
    uint4 pk;
    pk.x = (uint32_t)out[0] | ((uint32_t)out[1] << 16);
    pk.y = (uint32_t)out[2] | ((uint32_t)out[3] << 16);
    pk.z = (uint32_t)out[4] | ((uint32_t)out[5] << 16);
    pk.w = (uint32_t)out[6] | ((uint32_t)out[7] << 16);
    reinterpret_cast<uint4*>(h0 + (size_t)i * H)[t & 7] = pk;
  }
}

// dlambda = lam_prev + s (h3 w_out + b_out) (and y = h3 w_out + b_out), one row per thread; the
// 64 products are summed pairwise, c + 32, then c + 16, ... (the order of a 64-lane xor
// butterfly seen from lane 0)
__global__ void __launch_bounds__(kBlock)
k_decoder_predict(int n, const uint16_t* __restrict__ h3, const float* __restrict__ P,
                  const float* __restrict__ lp, const float* __restrict__ st,
                  float* __restrict__ y, float* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float p[H];
  const uint4* row = reinterpret_cast<const uint4*>(h3 + (size_t)i * H);
#pragma unroll
  for (int q = 0; q < H / 8; ++q) {
    const uint4 v = row[q];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      p[8 * q + 2 * e] = bf2f((uint16_t)(w[e] & 0xFFFFu)) * P[off_wout() + 8 * q + 2 * e];
      p[8 * q + 2 * e + 1] = bf2f((uint16_t)(w[e] >> 16)) * P[off_wout() + 8 * q + 2 * e + 1];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < off; ++c) p[c] = p[c] + p[c + off];
  const float yi = p[0] + P[off_bout()];
  if (y) y[i] = yi;
  if (out) out[i] = (lp ? lp[i] : 0.0f) + st[0] * yi;
}

// Decoder and top-layer ReLU backward: dz3[i][c] = dy_i w_out[c] where h3 > 0, else 0 (BF16),
// eight columns per thread
__global__ void __launch_bounds__(kBlock)
k_decoder_relu_bwd(int n, const float* __restrict__ dy, const float* __restrict__ P,
                   const uint16_t* __restrict__ h3, uint16_t* __restrict__ dz) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int i = t >> 3, c = 8 * (t & 7);
  if (i >= n) return;
  const float d = dy[i];
  const uint4 hv = *reinterpret_cast<const uint4*>(h3 + (size_t)i * H + c);
  const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
  uint32_t o[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float lo = bf2f((uint16_t)hw[e]) > 0.0f ? d * P[off_wout() + c + 2 * e] : 0.0f;
    const float hi = bf2f((uint16_t)(hw[e] >> 16)) > 0.0f ? d * P[off_wout() + c + 2 * e + 1] : 0.0f;
    o[e] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
  }
  *reinterpret_cast<uint4*>(dz + (size_t)i * H + c) = make_uint4(o[0], o[1], o[2], o[3]);
}

// column sums over rows (bias / decoder gradients), block partials: P[b][c]
__global__ void __launch_bounds__(kBlock)
k_colsum_partial(int n, int width, const uint16_t* __restrict__ Zbf, const float* __restrict__ Zf,
                 const float* __restrict__ rowscale, const uint16_t* __restrict__ Sbf,
                 float* __restrict__ partial) {
  // out[c] = sum_i (Z[i][c] or rowscale_i * S[i][c]) over the block's kColRows rows: G = 256 /
  // width row groups (group g takes rows g, g + G, ...), group totals added in group order
  __shared__ float sh[kBlock];
  const int G = kBlock / width;
  const int c = threadIdx.x % width, g = threadIdx.x / width;
  const int r0 = blockIdx.x * kColRows, r1 = min(n, r0 + kColRows);
  float s = 0.0f;
  if (g < G)
    for (int i = r0 + g; i < r1; i += G) {
      float v;
      if (Zbf) v = bf2f(Zbf[(size_t)i * width + c]);
      else if (Zf) v = Zf[(size_t)i * width + c];
      else v = rowscale[i] * bf2f(Sbf[(size_t)i * H + c]);
      s += v;
    }
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < width) {
    float t = 0.0f;
    for (int q = 0; q < G; ++q) t += sh[q * width + threadIdx.x];
    partial[(size_t)blockIdx.x * width + threadIdx.x] = t;
  }
}

// encoder weight gradient partials: P[b][k][c] = sum_i x[i][k] dz0[i][c] over the block's
// kColRows rows: thread (g, c) takes column c of rows g, g + 4, ... for all eight k (the row's
// features as two float4 loads), then the four groups are added in group order through LDS
__global__ void __launch_bounds__(kBlock)
k_encoder_wgrad(int n, const float* __restrict__ x, const uint16_t* __restrict__ dz0,
                float* __restrict__ partial) {
  __shared__ float sh[F][kBlock];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int r0 = blockIdx.x * kColRows, r1 = min(n, r0 + kColRows);
  float acc[F] = {};
  for (int i = r0 + g; i < r1; i += 4) {
    const float dz = bf2f(dz0[(size_t)i * H + c]);
    const float4 xa = reinterpret_cast<const float4*>(x)[2 * (size_t)i];
    const float4 xb = reinterpret_cast<const float4*>(x)[2 * (size_t)i + 1];
    const float xi[F] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
#pragma unroll
    for (int k = 0; k < F; ++k) acc[k] = fmaf(xi[k], dz, acc[k]);
  }
#pragma unroll
  for (int k = 0; k < F; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int o = threadIdx.x; o < F * H; o += kBlock) {
    const int k = o / H, cc = o % H;
    float t = 0.0f;
    for (int q = 0; q < 4; ++q) t += sh[k][64 * q + cc];
    partial[(size_t)blockIdx.x * F * H + o] = t;
  }
}

// Adam on the FP32 parameters, then BF16 copies of the message-passing weights:
// wt[l][0|1][out][in] (forward operand) and w[l][0|1][in][out] (backward operand).
__global__ void __launch_bounds__(kBlock)
k_adam(int np, float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= np) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  P[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(kBlock)
k_pack_bf16(const float* __restrict__ P, uint16_t* __restrict__ wt, uint16_t* __restrict__ w) {
  const int e = blockIdx.x * kBlock + threadIdx.x;   // over NL * 2 * H * H
  if (e >= NL * 2 * H * H) return;
  const int l = e / (2 * H * H);
  const int rem = e % (2 * H * H);
  const int half = rem / (H * H);
  const int in = (rem % (H * H)) / H, out = rem % H;
  const float val = P[off_wl(l) + (half * H + in) * H + out];
  wt[(size_t)(l * 2 + half) * H * H + out * H + in] = f2bf(val);
  w[(size_t)(l * 2 + half) * H * H + in * H + out] = f2bf(val);
}

// ------------------------------------------------------------------------------------------
// Features. x[i] = [jac/s, lp/s, C/sC, std log alpha, std log d, vol/mean - 1, deg/4, 1]
// with jac = b/A_ii, lp = previous lambda, s = rms(jac) (or rms(lp) when larger); the
// scalars come from FP64 sums passed in by the host-side reduction helper.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_feature_sums(int n, const float* __restrict__ b, const float* __restrict__ dinv,
               const float* __restrict__ lp, const float* __restrict__ C,
               const float* __restrict__ alpha, const float* __restrict__ vol,
               double* __restrict__ partial) {
  // per block: [sum jac^2, sum lp^2, sum C^2, sum la, sum la^2, sum ld, sum ld^2, sum vol]
  __shared__ double sh[8][kBlock];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double jac = (double)b[i] * (double)dinv[i];
    const double l = lp ? (double)lp[i] : 0.0;
    const double la = log((double)alpha[i]);
    const double ld = dinv[i] > 0.0f ? -log((double)dinv[i]) : 0.0;
    acc[0] += jac * jac; acc[1] += l * l; acc[2] += (double)C[i] * C[i];
    acc[3] += la; acc[4] += la * la; acc[5] += ld; acc[6] += ld * ld; acc[7] += vol[i];
  }
  for (int q = 0; q < 8; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 8; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 8) partial[blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

// One block of 256: strided per-thread sums of the block partials, then a fixed-order tree
// (deterministic).
__global__ void __launch_bounds__(256)
k_feature_stats(int n, int nblk, const double* __restrict__ partial, float* __restrict__ stats) {
  // stats: [s, sC, mean la, std la, mean ld, std ld, mean vol, unused]
  __shared__ double sh[8][256];
  const int tid = threadIdx.x;
  double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int bk = tid; bk < nblk; bk += 256)
    for (int q = 0; q < 8; ++q) t[q] += partial[(size_t)bk * 8 + q];
  for (int q = 0; q < 8; ++q) sh[q][tid] = t[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < 8; ++q) sh[q][tid] += sh[q][tid + w];
    __syncthreads();
  }
  if (tid != 0) return;
  for (int q = 0; q < 8; ++q) t[q] = sh[q][0];
  const double nn = (double)(n > 0 ? n : 1);
  const double sj = sqrt(t[0] / nn), sl = sqrt(t[1] / nn);
  const double s = fmax(fmax(sj, sl), 1e-30);
  const double ma = t[3] / nn, md = t[5] / nn;
  stats[0] = (float)s;
  stats[1] = (float)fmax(sqrt(t[2] / nn), 1e-30);
  stats[2] = (float)ma;
  stats[3] = (float)fmax(sqrt(fmax(t[4] / nn - ma * ma, 0.0)), 1e-6);
  stats[4] = (float)md;
  stats[5] = (float)fmax(sqrt(fmax(t[6] / nn - md * md, 0.0)), 1e-6);
  stats[6] = (float)fmax(t[7] / nn, 1e-30);
  stats[7] = 0.0f;
}

__global__ void __launch_bounds__(kBlock)
k_features(int n, const float* __restrict__ b, const float* __restrict__ dinv,
           const float* __restrict__ lp, const float* __restrict__ C,
           const float* __restrict__ alpha, const float* __restrict__ vol,
           const int32_t* __restrict__ nbr, const float* __restrict__ st,
           float* __restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float s = st[0];
  int d = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) d += nbr[4 * (size_t)i + q] >= 0;
  const float la = logf(alpha[i]);
  const float ld = dinv[i] > 0.0f ? -logf(dinv[i]) : st[4];
  float* o = x + (size_t)i * F;
  o[0] = b[i] * dinv[i] / s;
  o[1] = lp ? lp[i] / s : 0.0f;
  o[2] = C[i] / st[1];
  o[3] = (la - st[2]) / st[3];
  o[4] = (ld - st[4]) / st[5];
  o[5] = vol[i] / st[6] - 1.0f;
  o[6] = 0.25f * (float)d;
  o[7] = 1.0f;
}

// squared-error loss on the normalised output: dy = 2 (y - t) / n with t = (dlambda* - lp) / s,
// and per-block partials of the squared error
__global__ void __launch_bounds__(kBlock)
k_loss_mse(int n, const float* __restrict__ y, const float* __restrict__ target,
           const float* __restrict__ lp, const float* __restrict__ st, float* __restrict__ dy,
           double* __restrict__ partial) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float t = (target[i] - (lp ? lp[i] : 0.0f)) / st[0];
    const float e = y[i] - t;
    dy[i] = 2.0f * e / (float)n;
    acc += (double)e * (double)e;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// r = A x0 - b in place over ax, partials of r.r and b.b (2 per block)
__global__ void __launch_bounds__(kBlock)
k_resid_norms(int n, float* __restrict__ ax, const float* __restrict__ b,
              double* __restrict__ partial) {
  __shared__ double sh[2][kBlock];
  double rr = 0.0, bb = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float r = ax[i] - b[i];
    ax[i] = r;
    rr += (double)r * r;
    bb += (double)b[i] * b[i];
  }
  sh[0][threadIdx.x] = rr;
  sh[1][threadIdx.x] = bb;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = sh[0][0]; partial[2 * blockIdx.x + 1] = sh[1][0]; }
}

// scal[0] = r.r / b.b (the loss), scal[1] = 2 s / b.b (d loss / d y = scal[1] A r), from the
// (r.r, b.b) block pairs; one wavefront: lane l sums the pairs l, l + 64, ... in order, then an
// xor butterfly (deterministic, the same total in every lane). With den_io and den_ready the
// denominator b.b is read from den_io instead of the pairs' second entries; with den_io and
// !den_ready it is also stored there (the energy loss's t.A t, constant per training pair).
__global__ void __launch_bounds__(64)
k_resid_fin(int nblk, const double* __restrict__ partial, const float* __restrict__ st,
            double* __restrict__ scal, double* __restrict__ loss, double* __restrict__ den_io,
            int den_ready) {
  const int l = threadIdx.x;
  double rr = 0.0, bb = 0.0;
  for (int b = l; b < nblk; b += 64) {
    rr += partial[2 * b];
    if (!den_ready) bb += partial[2 * b + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rr += __shfl_xor(rr, off, 64);
    bb += __shfl_xor(bb, off, 64);
  }
  if (l != 0) return;
  if (den_ready) bb = *den_io;
  else if (den_io) *den_io = bb;
  const double inv = bb > 0.0 ? 1.0 / bb : 0.0;
  scal[0] = rr * inv;
  scal[1] = 2.0 * (double)st[0] * inv;
  if (loss) *loss = rr * inv;
}

// e = dlam - target (into e)
__global__ void __launch_bounds__(kBlock)
k_energy_err(int n, const float* __restrict__ dlam, const float* __restrict__ target,
             float* __restrict__ e) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) e[i] = dlam[i] - target[i];
}

// partial sums of a.b (one per block) into partial[slot + 2 * block]
__global__ void __launch_bounds__(kBlock)
k_pair_dot(int n, const float* __restrict__ a, const float* __restrict__ b,
           double* __restrict__ partial, int slot) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    acc += (double)a[i] * b[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[2 * blockIdx.x + slot] = sh[0];
}

__global__ void __launch_bounds__(kBlock)
k_resid_dy(int n, const float* __restrict__ ar, const double* __restrict__ scal,
           float* __restrict__ dy) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dy[i] = (float)(scal[1] * (double)ar[i]);
}

__global__ void k_sum_double(int nblk, const double* __restrict__ partial, double* __restrict__ out,
                             double scale) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[b];
  *out = s * scale;
}

}  // namespace mgpbd_gnn

using namespace mgpbd_gnn;

static inline int nb(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

struct GnnWorkspace {
  // activations: h[0..NL] and m[1..NL] BF16 n x 64; y n; dh FP32 n x 64; dz, dm BF16 n x 64
  uint16_t* h[NL + 1];
  uint16_t* m[NL + 1];
  float* y;
  float* dh;
  uint16_t* dz;
  uint16_t* dmb;
  float* partial;   // max(blocks) * 64 * 64
  float* partial2;  // the second weight gradient's block partials (same size)
  float* partialb;  // the bias gradient's block partials, blocks * 64
  double* dpartial;
  double* scal;
};

// Workspace layouts. Both start with the FP64 block partials and scalars (the features use
// only those), so a training workspace also serves the feature pass. Training keeps every
// layer's activations for the backward pass; inference ping-pongs two activation buffers.
static void carve(void* ws, int n, GnnWorkspace& W) {
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) / 256 * 256; return (void*)q; };
  W.dpartial = (double*)take((size_t)std::max(nb(n), 1) * 8 * 8);
  W.scal = (double*)take(64);
  const size_t act = (size_t)n * H * sizeof(uint16_t);
  for (int l = 0; l <= NL; ++l) W.h[l] = (uint16_t*)take(act);
  W.m[0] = nullptr;
  for (int l = 1; l <= NL; ++l) W.m[l] = (uint16_t*)take(act);
  W.y = (float*)take((size_t)n * 4);
  W.dh = (float*)take((size_t)n * H * 4);
  W.dz = (uint16_t*)take(act);
  W.dmb = (uint16_t*)take(act);
  const int nblk = (n + kWgradRows - 1) / kWgradRows;
  W.partial = (float*)take((size_t)std::max(nblk, 1) * H * H * 4);
  W.partial2 = (float*)take((size_t)std::max(nblk, 1) * H * H * 4);
  W.partialb = (float*)take((size_t)std::max(nblk, 1) * H * 4);
}

static void carve_infer(void* ws, int n, GnnWorkspace& W) {
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) / 256 * 256; return (void*)q; };
  W.dpartial = (double*)take((size_t)std::max(nb(n), 1) * 8 * 8);
  W.scal = (double*)take(64);
  const size_t act = (size_t)n * H * sizeof(uint16_t);
  W.h[0] = (uint16_t*)take(act);
  W.h[1] = (uint16_t*)take(act);
  for (int l = 2; l <= NL; ++l) W.h[l] = W.h[l & 1];
  W.y = (float*)take((size_t)n * 4);
}

static size_t infer_workspace_bytes(int n) {
  const size_t act = ((size_t)n * H * 2 + 255) / 256 * 256;
  return (((size_t)std::max(nb(n), 1) * 64 + 255) / 256 * 256) + 256 + 2 * act +
         ((size_t)n * 4 + 255) / 256 * 256;
}

static size_t workspace_bytes(int n) {
  const size_t act = ((size_t)n * H * 2 + 255) / 256 * 256;
  const size_t f4 = ((size_t)n * 4 + 255) / 256 * 256;
  const size_t fh = ((size_t)n * H * 4 + 255) / 256 * 256;
  const int nblk = (n + kWgradRows - 1) / kWgradRows;
  return act * (2 * NL + 1) + f4 + fh + 2 * act +
         2 * (((size_t)std::max(nblk, 1) * H * H * 4 + 255) / 256 * 256) +
         (((size_t)std::max(nblk, 1) * H * 4 + 255) / 256 * 256) +
         (((size_t)std::max(nb(n), 1) * 64 + 255) / 256 * 256) + 256;
}

extern "C" {

int mgpbd_gnn_param_count(void) { return n_params(); }
int64_t mgpbd_gnn_workspace_bytes(int n) { return n < 0 ? -1 : (int64_t)workspace_bytes(n); }
int64_t mgpbd_gnn_infer_workspace_bytes(int n) {
  return n < 0 ? -1 : (int64_t)infer_workspace_bytes(n);
}

int mgpbd_gnn_features(mgpbd_ctx* ctx, int n, const float* b, const float* dinv,
                       const float* lam_prev, const float* C, const float* alpha,
                       const float* vol, const int32_t* nbr, float* x, float* stats, void* ws) {
  if (!ctx || n <= 0 || !b || !dinv || !C || !alpha || !vol || !nbr || !x || !stats || !ws)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_features: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  const int g = std::min(nb(n), 1024);
  k_feature_sums<<<g, kBlock, 0, s>>>(n, b, dinv, lam_prev, C, alpha, vol, W.dpartial);
  k_feature_stats<<<1, 256, 0, s>>>(n, g, W.dpartial, stats);
  k_features<<<nb(n), kBlock, 0, s>>>(n, b, dinv, lam_prev, C, alpha, vol, nbr, stats, x);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_features");
  return 0;
}

int mgpbd_gnn_forward(mgpbd_ctx* ctx, int n, const int32_t* nbr, const float* x,
                      const float* stats, const float* lam_prev, const float* params,
                      const uint16_t* wt, void* ws, float* dlam_out) {
  if (!ctx || n <= 0 || !nbr || !x || !stats || !params || !wt || !ws)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_forward: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  // the inference kernels, keeping every layer's input and neighbour mean for the backward
  // pass
  k_encoder8<<<std::min(nb((int64_t)n * 8), kEncoderBlocks), kBlock, 0, s>>>(n, x, params,
                                                                             W.h[0]);
  const int gg = std::min((n + 32 * kTileWaves - 1) / (32 * kTileWaves), kLayerBlocks);
  for (int l = 0; l < NL; ++l) {
    const uint16_t* w_self = wt + (size_t)(2 * l) * H * H;
    const uint16_t* w_nbr = wt + (size_t)(2 * l + 1) * H * H;
    k_layer_mean<<<gg, kBlock, 0, s>>>(n, nbr, W.h[l], w_self, w_nbr, params + off_bl(l),
                                       W.h[l + 1], W.m[l + 1], FusedDecoder{});
  }
  k_decoder_predict<<<nb(n), kBlock, 0, s>>>(n, W.h[NL], params, lam_prev, stats, W.y,
                                             dlam_out);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_forward");
  return 0;
}

int mgpbd_gnn_infer(mgpbd_ctx* ctx, int n, const int32_t* nbr, const float* x,
                    const float* stats, const float* lam_prev, const float* params,
                    const uint16_t* wt, void* ws, float* dlam_out) {
  if (!ctx || n <= 0 || !nbr || !x || !stats || !params || !wt || !ws || !dlam_out)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_infer: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve_infer(ws, n, W);   // layer l reads h[l & 1] and writes h[(l + 1) & 1]
  k_encoder8<<<std::min(nb((int64_t)n * 8), kEncoderBlocks), kBlock, 0, s>>>(n, x, params,
                                                                             W.h[0]);
  const int gg = std::min((n + 32 * kTileWaves - 1) / (32 * kTileWaves), kLayerBlocks);
  for (int l = 0; l < NL; ++l) {
    const uint16_t* w_self = wt + (size_t)(2 * l) * H * H;
    const uint16_t* w_nbr = wt + (size_t)(2 * l + 1) * H * H;
    // the last layer feeds the decoder from LDS (its rows never reach memory)
    const FusedDecoder dec = l == NL - 1 ? FusedDecoder{params, lam_prev, stats, W.y, dlam_out}
                                         : FusedDecoder{};
    k_layer_mean<<<gg, kBlock, 0, s>>>(n, nbr, W.h[l], w_self, w_nbr, params + off_bl(l),
                                       W.h[l + 1], nullptr, dec);
  }
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_infer");
  return 0;
}

int mgpbd_gnn_loss_mse(mgpbd_ctx* ctx, int n, const float* stats, const float* lam_prev,
                       const float* target, void* ws, float* dy, double* loss_out) {
  if (!ctx || n <= 0 || !stats || !target || !ws || !dy)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_loss_mse: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  const int gl = std::min(nb(n), 1024);
  k_loss_mse<<<gl, kBlock, 0, s>>>(n, W.y, target, lam_prev, stats, dy, W.dpartial);
  if (loss_out) k_sum_double<<<1, 64, 0, s>>>(gl, W.dpartial, loss_out, 1.0 / n);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_loss_mse");
  return 0;
}

int mgpbd_gnn_loss_residual(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr,
                            const int32_t* sell_idx, const float* sell_val, const float* b,
                            const float* dlam, const float* stats, float* tmp, void* ws,
                            float* dy, double* loss_out) {
  if (!ctx || n <= 0 || !slice_ptr || !sell_idx || !sell_val || !b || !dlam || !stats || !tmp ||
      !ws || !dy)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_loss_residual: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  float* r = tmp;
  float* ar = tmp + n;
  if (int e = mgpbd_spmv(ctx, n, slice_ptr, sell_idx, sell_val, dlam, r)) return e;
  const int gl = std::min(nb(n), 1024);
  k_resid_norms<<<gl, kBlock, 0, s>>>(n, r, b, W.dpartial);
  k_resid_fin<<<1, 64, 0, s>>>(gl, W.dpartial, stats, W.scal, loss_out, nullptr, 0);
  if (int e = mgpbd_spmv(ctx, n, slice_ptr, sell_idx, sell_val, r, ar)) return e;
  k_resid_dy<<<nb(n), kBlock, 0, s>>>(n, ar, W.scal, dy);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_loss_residual");
  return 0;
}

int mgpbd_gnn_loss_energy(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr,
                          const int32_t* sell_idx, const float* sell_val, const float* target,
                          const float* dlam, const float* stats, float* tmp, void* ws, float* dy,
                          double* loss_out, double* tat_io, int tat_ready) {
  if (!ctx || n <= 0 || !slice_ptr || !sell_idx || !sell_val || !target || !dlam || !stats ||
      !tmp || !ws || !dy)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_loss_energy: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  float* e = tmp;
  float* ae = tmp + n;
  const int gl = std::min(nb(n), 1024);
  k_energy_err<<<nb(n), kBlock, 0, s>>>(n, dlam, target, e);
  if (int r = mgpbd_spmv(ctx, n, slice_ptr, sell_idx, sell_val, e, ae)) return r;
  k_pair_dot<<<gl, kBlock, 0, s>>>(n, e, ae, W.dpartial, 0);          // e.Ae
  const bool cached = tat_io && tat_ready;
  if (!cached) {   // t.At, constant per training pair: once when the caller keeps it
    if (int r = mgpbd_spmv(ctx, n, slice_ptr, sell_idx, sell_val, target, e)) return r;
    k_pair_dot<<<gl, kBlock, 0, s>>>(n, target, e, W.dpartial, 1);   // t.At
  }
  // same finalisation as the residual loss: scal = {e.Ae / t.At, 2 s / t.At}
  k_resid_fin<<<1, 64, 0, s>>>(gl, W.dpartial, stats, W.scal, loss_out, tat_io, cached ? 1 : 0);
  k_resid_dy<<<nb(n), kBlock, 0, s>>>(n, ae, W.scal, dy);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_loss_energy");
  return 0;
}

int mgpbd_gnn_backward(mgpbd_ctx* ctx, int n, const int32_t* nbr, const float* x,
                       const float* params, const uint16_t* w, const float* dy, void* ws,
                       float* grad) {
  if (!ctx || n <= 0 || !nbr || !x || !params || !w || !dy || !ws || !grad)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_backward: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  GnnWorkspace W;
  carve(ws, n, W);
  const int nblk = (n + kWgradRows - 1) / kWgradRows;
  const int ncs = (n + kColRows - 1) / kColRows;
  // decoder: d w_out = sum_i dy_i h3_i, d b_out = sum_i dy_i, dh3 = dy w_out^T
  k_colsum_partial<<<ncs, kBlock, 0, s>>>(n, H, nullptr, nullptr, dy, W.h[NL], W.partial);
  sum_partials(s, ncs, H, W.partial, grad + off_wout());
  k_colsum_partial<<<ncs, kBlock, 0, s>>>(n, 1, nullptr, dy, nullptr, nullptr, W.partial);
  sum_partials(s, ncs, 1, W.partial, grad + off_bout());
  // dZ of the top layer = (dy w_out^T) * (h3 > 0)
  k_decoder_relu_bwd<<<nb((int64_t)n * 8), kBlock, 0, s>>>(n, dy, params, W.h[NL], W.dz);
  const int gg = (n + 127) / 128;
  for (int l = NL - 1; l >= 0; --l) {
    // weight and bias gradients: one pass over dZ for W_self, W_nbr and b
    k_wgrad<<<nblk, kBlock, 0, s>>>(n, W.h[l], W.dz, W.partial, W.m[l + 1], W.partial2,
                                    W.partialb);
    sum_partials(s, nblk, H * H, W.partial, grad + off_wl(l));
    sum_partials(s, nblk, H * H, W.partial2, grad + off_wl(l) + H * H);
    sum_partials(s, nblk, H, W.partialb, grad + off_bl(l));
    // input gradients: dh_l = dZ W_self^T + mean-backward(dZ W_nbr^T), then the ReLU of h_l:
    // dZ of the layer below (of the encoder for l = 0) = dh_l * (h_l > 0)
    const uint16_t* w_self = w + (size_t)(2 * l) * H * H;
    const uint16_t* w_nbr = w + (size_t)(2 * l + 1) * H * H;
    k_gemm64_dual<<<gg, kBlock, 0, s>>>(n, W.dz, w_nbr, w_self, W.dmb, W.dh);
    k_mean_nbr_relu_bwd<<<nb((int64_t)n * 8), kBlock, 0, s>>>(n, nbr, W.dmb, W.dh, W.h[l], W.dz);
  }
  // encoder: dW_in = x^T dZ0; db_in = sum dZ0
  k_encoder_wgrad<<<ncs, kBlock, 0, s>>>(n, x, W.dz, W.partial);
  sum_partials(s, ncs, F * H, W.partial, grad + off_win());
  k_colsum_partial<<<ncs, kBlock, 0, s>>>(n, H, W.dz, nullptr, nullptr, nullptr, W.partial);
  sum_partials(s, ncs, H, W.partial, grad + off_bin());
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_backward");
  return 0;
}

int mgpbd_gnn_adam(mgpbd_ctx* ctx, float* params, const float* grad, float* m, float* v,
                   float lr, float beta1, float beta2, float eps, int step, uint16_t* wt,
                   uint16_t* w) {
  if (!ctx || !params || !grad || !m || !v || step < 1)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_adam: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  k_adam<<<nb(n_params()), kBlock, 0, s>>>(n_params(), params, grad, m, v, lr, beta1, beta2, eps, c1, c2);
  if (wt && w) k_pack_bf16<<<nb(NL * 2 * H * H), kBlock, 0, s>>>(params, wt, w);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_adam");
  return 0;
}

int mgpbd_gnn_pack(mgpbd_ctx* ctx, const float* params, uint16_t* wt, uint16_t* w) {
  if (!ctx || !params || !wt || !w) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_pack: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  k_pack_bf16<<<nb(NL * 2 * H * H), kBlock, 0, s>>>(params, wt, w);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_pack");
  return 0;
}

// Unit entry for tests: Y = act(A1 W1 + A2 W2 + bias) with the MFMA kernel (BF16 operands).
int mgpbd_gnn_gemm64(mgpbd_ctx* ctx, int n, const uint16_t* A1, const uint16_t* W1t,
                     const uint16_t* A2, const uint16_t* W2t, const float* bias, int relu,
                     float* Y) {
  if (!ctx || n <= 0 || !A1 || !W1t || !Y || (A2 && !W2t))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_gemm64: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  k_gemm64<<<(n + 127) / 128, kBlock, 0, s>>>(n, A1, W1t, A2, W2t, bias, relu, nullptr, Y);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_gemm64");
  return 0;
}

// Unit entry for tests: G[64][64] = A^T dZ (A, dZ BF16 n x 64) with the MFMA reduction.
int mgpbd_gnn_wgrad64(mgpbd_ctx* ctx, int n, const uint16_t* A, const uint16_t* dZ,
                      float* partial, float* G) {
  if (!ctx || n <= 0 || !A || !dZ || !partial || !G)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_gnn_wgrad64: invalid arguments");
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  const int nblk = (n + kWgradRows - 1) / kWgradRows;
  k_wgrad<<<nblk, kBlock, 0, s>>>(n, A, dZ, partial, nullptr, nullptr, nullptr);
  sum_partials(s, nblk, H * H, partial, G);
  MGPBD_LAUNCH_CHECK("mgpbd_gnn_wgrad64");
  return 0;
}

}  // extern "C"
