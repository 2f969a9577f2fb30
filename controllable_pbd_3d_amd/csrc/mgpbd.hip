// gfx950 kernels of the MGPBD global XPBD solve and their C-ABI launchers.
//
// Layout in HBM (one GPU):
//   positions / velocities      float[3 * n_verts]   (AoS, the caller's (n, 3) arrays)
//   t_i                         int32[4 * n_tets]    (loaded as int4)
//   Dm_inv                      float[9 * n_tets]    row-major 3x3
//   gradC                       float[12 * n_tets]   4 vertex gradients x 3
//   level matrices              SELL-64: per 64-row slice, entry k of row 64s+l at
//                               slice_ptr[s] + 64k + l; int32 columns, float values
//
// Arithmetic conventions shared with the CPU restatement in oracle/: products and sums are
// written out with explicit fmaf where the order matters, so the sparse kernels (SpMV,
// Jacobi, refill, restrict, prolong) and the A fill produce identical FP32 bits on both
// sides for identical inputs. Reductions accumulate in FP64 with a fixed tree order
// (deterministic run to run).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "mgpbd_internal.h"
#include "svd3.h"

namespace mgpbd {

constexpr int kBlock = 256;           // 4 wavefronts of 64
constexpr int kMaxRedBlocks = 2048;   // grid cap for grid-stride reduction kernels
constexpr int kPartialStride = 65536; // partials per reduction slot (full-grid kernels)
constexpr int kPartialSlots = 5;       // reduction slots of a context (slots 0 .. 4 are used)

// ------------------------------------------------------------------------------------------
// Deterministic FP64 reductions in two phases. Phase 1: every block of the producing kernel
// stores its block total with a plain store. Phase 2: a one-block kernel sums the per-block
// totals in a fixed order and runs the scalar epilogue. The kernel boundary orders the two
// phases, so no block pays a device-scope fence (on gfx950 an agent-scope release writes
// back the XCD's L2, which costs ~20 ns per block) and the result is the same every run.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Returns the block total in thread 0 (other threads: unspecified).
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kBlock / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return s;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8), each XCD with
// its own L2. Remapping block b to a contiguous range per XCD keeps neighbouring work (rows
// sharing cache lines of their gathered inputs) on one L2; a bijection on [0, G).
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int x = b & 7, j = b >> 3, q = G >> 3, rem = G & 7;
  return x * q + min(x, rem) + j;
}

struct RedWS {
  double* partials;       // kPartialStride per slot, kPartialSlots slots
};

// slot < kPartialSlots (every call site passes a constant)
__device__ __forceinline__ void store_partial(RedWS ws, int slot, double v) {
  if (threadIdx.x == 0) ws.partials[(size_t)slot * kPartialStride + blockIdx.x] = v;
}

// Phase 2 helper (one block): totals of slots first .. first + K - 1 over `nparts` block
// partials (thread 0). Each thread keeps four independent partial sums per slot, so up to
// 4 K loads are in flight at once instead of one dependent load per step (the partials come
// from other XCDs' L2 write-backs: each round trip is long); the order is fixed.
template <int K>
__device__ __forceinline__ void sum_partials_k(RedWS ws, int first, int nparts,
                                               double (&out)[K]) {
  const double* p = ws.partials + (size_t)first * kPartialStride;
  double a[K][4];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k][0] = a[k][1] = a[k][2] = a[k][3] = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * kBlock < nparts; i += 4 * kBlock) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) a[k][j] += p[(size_t)k * kPartialStride + i + j * kBlock];
  }
  for (; i < nparts; i += kBlock)
#pragma unroll
    for (int k = 0; k < K; ++k) a[k][0] += p[(size_t)k * kPartialStride + i];
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = block_sum((a[k][0] + a[k][1]) + (a[k][2] + a[k][3]));
}

__device__ __forceinline__ double sum_partials(RedWS ws, int slot, int nparts) {
  double t[1];
  sum_partials_k<1>(ws, slot, nparts, t);
  return t[0];
}

// out[s] = total of slot s, for s in [first, nslots)
__global__ void __launch_bounds__(kBlock)
k_fin_sum(RedWS ws, int nparts, int first, int nslots, double* __restrict__ out) {
  for (int sl = first; sl < nslots; ++sl) {
    double t = sum_partials(ws, sl, nparts);
    if (threadIdx.x == 0) out[sl] = t;
  }
}

// ------------------------------------------------------------------------------------------
// SELL-64x4: entry k of row r at slice_ptr[r/64] + 256*(k/4) + 4*(r%64) + (k%4).
// ------------------------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned short v4u16 __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int sell_base(const int* __restrict__ slice_ptr, int r) {
  return slice_ptr[r >> 6] + 4 * (r & 63);
}
__device__ __forceinline__ int sell_width(const int* __restrict__ slice_ptr, int r) {
  return (slice_ptr[(r >> 6) + 1] - slice_ptr[r >> 6]) >> 6;
}

// ------------------------------------------------------------------------------------------
// Per-tet kernels
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void load3(const float* __restrict__ p, int v, float* o) {
  o[0] = p[3 * (size_t)v + 0];
  o[1] = p[3 * (size_t)v + 1];
  o[2] = p[3 * (size_t)v + 2];
}

// Rest shape in FP64 (rounded once to FP32): Dm^-1 = adj(Dm) / det, vol = |det| / 6,
// alpha = 1 / (mu dt^2 vol). With the volume constraint (lam_t or lam > 0 given, vol_rows)
// tet t owns constraints 2t (ARAP) and 2t+1 (volume, alpha = 1 / (lambda dt^2 vol)).
__global__ void __launch_bounds__(kBlock)
k_rest_shape(int n, const int4* __restrict__ t_i, const float* __restrict__ xr,
             const float* __restrict__ mu_t, double mu, const float* __restrict__ lam_t,
             double lam, int vol_rows, double dt, float* __restrict__ Dm_inv,
             float* __restrict__ vol_out, float* __restrict__ alpha) {
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  int4 tv = t_i[t];
  float r0[3], r1[3], r2[3], r3[3];
  load3(xr, tv.x, r0); load3(xr, tv.y, r1); load3(xr, tv.z, r2); load3(xr, tv.w, r3);
  double D[3][3];   // Dm = [r1-r0, r2-r0, r3-r0] as columns; D[row][col]
  for (int r = 0; r < 3; ++r) {
    D[r][0] = (double)r1[r] - (double)r0[r];
    D[r][1] = (double)r2[r] - (double)r0[r];
    D[r][2] = (double)r3[r] - (double)r0[r];
  }
  double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
  double c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
  double c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
  double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
  float* o = Dm_inv + 9 * (size_t)t;
  // inverse = adjugate / det, adj[i][j] = cofactor[j][i]
  o[0] = (float)(c00 / det);
  o[1] = (float)((D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det);
  o[2] = (float)((D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det);
  o[3] = (float)(c01 / det);
  o[4] = (float)((D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det);
  o[5] = (float)((D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det);
  o[6] = (float)(c02 / det);
  o[7] = (float)((D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det);
  o[8] = (float)((D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det);
  double vol = fabs(det) / 6.0;
  vol_out[t] = (float)vol;
  double m = mu_t ? (double)mu_t[t] : mu;
  if (vol_rows) {
    double l = lam_t ? (double)lam_t[t] : lam;
    alpha[2 * (size_t)t] = (float)(1.0 / (m * dt * dt * vol));
    alpha[2 * (size_t)t + 1] = (float)(1.0 / (l * dt * dt * vol));
  } else {
    alpha[t] = (float)(1.0 / (m * dt * dt * vol));
  }
}

// F = Ds Dm^-1 (FP64) for tet t at positions x (+ omega * dx in FP32 when dx != nullptr,
// the trial positions of the line search).
__device__ __forceinline__ void deformation_gradient(const int4 tv, const float* __restrict__ x,
                                                     const float* __restrict__ dx, float omega,
                                                     const float* __restrict__ Bf, double F[9]) {
  double p[4][3];
  int vs[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    float q[3];
    load3(x, vs[l], q);
    if (dx) {
      float d[3];
      load3(dx, vs[l], d);
#pragma unroll
      for (int r = 0; r < 3; ++r) q[r] = q[r] + omega * d[r];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) p[l][r] = (double)q[r];
  }
  double Ds[3][3], B[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) B[k] = (double)Bf[k];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    Ds[r][0] = p[1][r] - p[0][r];
    Ds[r][1] = p[2][r] - p[0][r];
    Ds[r][2] = p[3][r] - p[0][r];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      F[3 * r + c] = Ds[r][0] * B[c] + Ds[r][1] * B[3 + c] + Ds[r][2] * B[6 + c];
}

__device__ __forceinline__ double arap_C(const double s[3]) {
  double a = s[0] - 1.0, b = s[1] - 1.0, c = s[2] - 1.0;
  return sqrt(a * a + b * b + c * c);
}

// ARAP constraint C = sqrt(sum_i (s_i - 1)^2) over F's rotation-variant singular values and,
// with GRAD, dC/dF = U diag((s_i - 1) / C) V^T into G when (float)C > 1e-6 (else G untouched).
// For det F > 0 from the polar rotation R = U V^T (C = ||F - R||_F, dC/dF = (F - R) / C, a
// few Newton steps); inverted, collapsed or strongly compressed tets take the Jacobi SVD.
template <bool GRAD>
__device__ __forceinline__ double arap_constraint(const double F[9], double (&G)[3][3]) {
  double R[9];
  if (polar_rotation(F, R)) {
    double D[9], c2 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      D[k] = F[k] - R[k];
      c2 = fma(D[k], D[k], c2);
    }
    const double Cd = sqrt(c2);
    if (GRAD && (float)Cd > 1e-6f) {
#pragma unroll
      for (int k = 0; k < 9; ++k) G[k / 3][k % 3] = D[k] / Cd;
    }
    return Cd;
  }
  double U[3][3], s[3], V[3][3];
  svd3(F, U, s, V);
  const double Cd = arap_C(s);
  if (GRAD && (float)Cd > 1e-6f) {
    const double d[3] = {(s[0] - 1.0) / Cd, (s[1] - 1.0) / Cd, (s[2] - 1.0) / Cd};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        G[r][c] = U[0][r] * V[0][c] * d[0] + U[1][r] * V[1][c] * d[1] + U[2][r] * V[2][c] * d[2];
  }
  return Cd;
}

// Volume constraint det F - 1 and dC/dF = cof(F): column c of the cofactor matrix is the
// cross product of F's columns c+1 and c+2 (cyclic). G[r][c], returns det F - 1.
__device__ __forceinline__ double volume_C(const double F[9], double G[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    G[0][c] = F[3 + a] * F[6 + b] - F[6 + a] * F[3 + b];
    G[1][c] = F[6 + a] * F[0 + b] - F[0 + a] * F[6 + b];
    G[2][c] = F[0 + a] * F[3 + b] - F[3 + a] * F[0 + b];
  }
  const double det = F[0] * G[0][0] + F[3] * G[1][0] + F[6] * G[2][0];
  return det - 1.0;
}

// Vertex gradients g_k = G Dm^-T[:, k-1] (k = 1..3), g_0 = -(g_1 + g_2 + g_3), rounded to
// FP32 and stored as three float4 at gradC + 12 c.
__device__ __forceinline__ void store_gradient(const double G[3][3], const float Bf[9],
                                               float* __restrict__ gradC, size_t c) {
  double gd[4][3];
#pragma unroll
  for (int k = 1; k <= 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      gd[k][r] = G[r][0] * (double)Bf[3 * (k - 1)] + G[r][1] * (double)Bf[3 * (k - 1) + 1] +
                 G[r][2] * (double)Bf[3 * (k - 1) + 2];
#pragma unroll
  for (int r = 0; r < 3; ++r) gd[0][r] = -gd[1][r] - gd[2][r] - gd[3][r];
  float g[12];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) g[3 * k + r] = (float)gd[k][r];
  float4* go = reinterpret_cast<float4*>(gradC + 12 * c);
  go[0] = make_float4(g[0], g[1], g[2], g[3]);
  go[1] = make_float4(g[4], g[5], g[6], g[7]);
  go[2] = make_float4(g[8], g[9], g[10], g[11]);
}

// VOL: two constraints per tet, 2t the ARAP constraint and 2t+1 the volume constraint (its
// gradient is never zeroed); the per-constraint arrays hold 2 n entries.
template <bool VOL>
__global__ void __launch_bounds__(kBlock)
k_constraints(int n, const int4* __restrict__ t_i, const float* __restrict__ x,
              const float* __restrict__ Dm_inv, const float* __restrict__ alpha,
              const float* __restrict__ lambda, float* __restrict__ C_out,
              float* __restrict__ gradC, float* __restrict__ rhs, RedWS ws) {
  double acc = 0.0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
    const size_t c0 = VOL ? 2 * (size_t)t : (size_t)t;
    int4 tv = t_i[t];
    float Bf[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Bf[k] = Dm_inv[9 * (size_t)t + k];
    double F[9];
    deformation_gradient(tv, x, nullptr, 0.0f, Bf, F);
    double G[3][3];   // dC/dF
    const double Cd = arap_constraint<true>(F, G);
    float C = (float)Cd;
    float g[12];
    if (C > 1e-6f) {
      // g_k = G B^T[:, k-1] (k = 1..3), g_0 = -(g_1 + g_2 + g_3)
      double gd[4][3];
#pragma unroll
      for (int k = 1; k <= 3; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          gd[k][r] = G[r][0] * (double)Bf[3 * (k - 1)] + G[r][1] * (double)Bf[3 * (k - 1) + 1] +
                     G[r][2] * (double)Bf[3 * (k - 1) + 2];
#pragma unroll
      for (int r = 0; r < 3; ++r) gd[0][r] = -gd[1][r] - gd[2][r] - gd[3][r];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r) g[3 * k + r] = (float)gd[k][r];
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) g[k] = 0.0f;
    }
    C_out[c0] = C;
    float4* go = reinterpret_cast<float4*>(gradC + 12 * c0);
    go[0] = make_float4(g[0], g[1], g[2], g[3]);
    go[1] = make_float4(g[4], g[5], g[6], g[7]);
    go[2] = make_float4(g[8], g[9], g[10], g[11]);
    float b = -(C + alpha[c0] * lambda[c0]);
    rhs[c0] = b;
    acc += (double)b * (double)b;
    if (VOL) {
      double Gh[3][3];
      const float Ch = (float)volume_C(F, Gh);
      store_gradient(Gh, Bf, gradC, c0 + 1);
      C_out[c0 + 1] = Ch;
      const float bh = -(Ch + alpha[c0 + 1] * lambda[c0 + 1]);
      rhs[c0 + 1] = bh;
      acc += (double)bh * (double)bh;
    }
  }
  double bs = block_sum(acc);
  store_partial(ws, 0, bs);
}

template <bool VOL>
__global__ void __launch_bounds__(kBlock)
k_dual_norms(int n, const int4* __restrict__ t_i, const float* __restrict__ x,
             const float* __restrict__ dx, float omega, const float* __restrict__ Dm_inv,
             const float* __restrict__ alpha, const float* __restrict__ lambda,
             const float* __restrict__ C_cur, int with_current, RedWS ws) {
  double acc_cur = 0.0, acc_trial = 0.0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
    const size_t c0 = VOL ? 2 * (size_t)t : (size_t)t;
    float al = alpha[c0] * lambda[c0];
    if (with_current != 2) {   // 2: the current norm only
      int4 tv = t_i[t];
      float B[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) B[k] = Dm_inv[9 * (size_t)t + k];
      double F[9];
      deformation_gradient(tv, x, dx, omega, B, F);
      double G[3][3];
      float c = (float)arap_constraint<false>(F, G);
      float r = -(c + al);
      acc_trial += (double)r * (double)r;
      if (VOL) {
        double Gh[3][3];
        const float ch = (float)volume_C(F, Gh);
        const float rh = -(ch + alpha[c0 + 1] * lambda[c0 + 1]);
        acc_trial += (double)rh * (double)rh;
      }
    }
    if (with_current) {
      float rc = -(C_cur[c0] + al);
      acc_cur += (double)rc * (double)rc;
      if (VOL) {
        const float rh = -(C_cur[c0 + 1] + alpha[c0 + 1] * lambda[c0 + 1]);
        acc_cur += (double)rh * (double)rh;
      }
    }
  }
  double b0 = block_sum(acc_cur);
  double b1 = block_sum(acc_trial);
  store_partial(ws, 0, b0);
  store_partial(ws, 1, b1);
}

// Off-diagonal entry from tet j's vertex ids and gradient (already loaded).
__device__ __forceinline__ float a_offdiag(const int (&vi)[4], const float (&w)[4],
                                           const float (&gi)[4][3], int4 tj, float4 q0,
                                           float4 q1, float4 q2) {
  const int vj[4] = {tj.x, tj.y, tj.z, tj.w};
  const float gj[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
  float v = 0.0f;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (vi[a] == vj[b]) {
        float g[3] = {gj[3 * b], gj[3 * b + 1], gj[3 * b + 2]};
        v = fmaf(w[a], dot3(gi[a], g), v);
      }
  return v;
}

// One A entry (row i, column j): the diagonal sum_v w_v |g_iv|^2 + alpha_i, or the sum over
// the vertices i and j share of w_v g_iv . g_jv in the slot order of _intersect_tets.
__device__ __forceinline__ float a_entry(int i, int j, const int (&vi)[4], const float (&w)[4],
                                         const float (&gi)[4][3], float alpha_i,
                                         const int4* __restrict__ t_i,
                                         const float* __restrict__ gradC) {
  if (j == i) {
    float v = w[0] * dot3(gi[0], gi[0]);
    v = fmaf(w[1], dot3(gi[1], gi[1]), v);
    v = fmaf(w[2], dot3(gi[2], gi[2]), v);
    v = fmaf(w[3], dot3(gi[3], gi[3]), v);
    return v + alpha_i;
  }
  const int4 tj = t_i[j];
  // the 48 bytes of gradC_j in three 16-byte loads
  const float4* g4 = reinterpret_cast<const float4*>(gradC) + 3 * (size_t)j;
  return a_offdiag(vi, w, gi, tj, g4[0], g4[1], g4[2]);
}


// A values, one thread per row (tet) i, four entries per 16-byte store (padding stays 0).
__global__ void __launch_bounds__(kBlock)
k_fill_A(int n, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
         const int* __restrict__ row_len, const int4* __restrict__ t_i,
         const float* __restrict__ invm, const float* __restrict__ gradC,
         const float* __restrict__ alpha, float* __restrict__ val, float* __restrict__ dinv) {
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int base = sell_base(slice_ptr, i);
  const int len = row_len[i];
  const int4 ti = t_i[i];
  const int vi[4] = {ti.x, ti.y, ti.z, ti.w};
  float w[4], gi[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    w[a] = invm[vi[a]];
#pragma unroll
    for (int r = 0; r < 3; ++r) gi[a][r] = gradC[12 * (size_t)i + 3 * a + r];
  }
  const float al = alpha[i];
  float diag = 0.0f;
  const float4* g4 = reinterpret_cast<const float4*>(gradC);
  for (int k0 = 0; k0 < len; k0 += 4) {
    const int p = base + 64 * k0;   // 256 * (k0 / 4)
    const int4 c = *reinterpret_cast<const int4*>(idx + p);
    const int cj[4] = {c.x, c.y, c.z, c.w};
    // the four neighbours' vertex ids and gradients, all loads issued before any arithmetic
    // (padding slots and the diagonal load row i's own, unused)
    int4 tj[4];
    float4 q0[4], q1[4], q2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = k0 + q < len ? cj[q] : i;
      tj[q] = t_i[j];
      q0[q] = g4[3 * (size_t)j];
      q1[q] = g4[3 * (size_t)j + 1];
      q2[q] = g4[3 * (size_t)j + 2];
    }
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = 0.0f;
      if (k0 + q < len) {
        if (cj[q] == i) {
          v[q] = a_entry(i, i, vi, w, gi, al, t_i, gradC);
          diag = v[q];
        } else {
          v[q] = a_offdiag(vi, w, gi, tj[q], q0[q], q1[q], q2[q]);
        }
      }
    }
    *reinterpret_cast<float4*>(val + p) = make_float4(v[0], v[1], v[2], v[3]);
  }
  dinv[i] = fabsf(diag) > 1e-30f ? 1.0f / diag : 0.0f;
}

// Shared-vertex descriptors of the level-0 pattern, one 16-bit word per SELL slot: bits 12-13
// the number of vertices row i shares with column j (1-3), and per shared vertex, in ascending
// local slot a of tet i, 4 bits (a | b << 2) with b the vertex's local slot in tet j;
// kDescDiag marks the diagonal, 0 padding. Rows that share all four vertices (the ARAP and
// volume constraints of one tet) are kDescAll with, per local slot a of row i, 2 bits at 2a:
// the slot b of the same vertex in row j. The pattern and t_i fix them: built once.
constexpr uint16_t kDescDiag = 0x8000;
constexpr uint16_t kDescAll = 0x4000;

__global__ void __launch_bounds__(kBlock)
k_fill_desc(int n, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
            const int* __restrict__ row_len, const int4* __restrict__ t_i,
            uint16_t* __restrict__ desc) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= (n + 63) >> 6 || i >= n) return;
  const int base = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const int len = row_len[i];
  const int4 ti = t_i[i];
  const int vi[4] = {ti.x, ti.y, ti.z, ti.w};
  for (int k = 0; k < w; ++k) {
    const int p = base + 64 * (k & ~3) + (k & 3);
    uint16_t d = 0;
    if (k < len) {
      const int j = idx[p];
      if (j == i) {
        d = kDescDiag;
      } else {
        const int4 tj = t_i[j];
        const int vj[4] = {tj.x, tj.y, tj.z, tj.w};
        int cnt = 0, code = 0, perm = 0;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b)
            if (vi[a] == vj[b]) {
              if (cnt < 3) code |= (a | (b << 2)) << (4 * cnt);
              perm |= b << (2 * a);
              ++cnt;
            }
        d = cnt == 4 ? (uint16_t)(kDescAll | perm) : (uint16_t)((cnt << 12) | code);
      }
    }
    desc[p] = d;
  }
}

// The A fill from the descriptors: per off-diagonal slot only tet j's gradients at the shared
// vertices are gathered (12 bytes each instead of j's 16-byte vertex ids and 48-byte
// gradient); the products are summed in the same order as k_fill_A (ascending local slot of
// tet i), so the values are identical.
__global__ void __launch_bounds__(kBlock)
k_fill_A_desc(int n, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
              const int* __restrict__ row_len, const uint16_t* __restrict__ desc,
              const int4* __restrict__ t_i, const float* __restrict__ invm,
              const float* __restrict__ gradC, const float* __restrict__ alpha,
              float* __restrict__ val, float* __restrict__ dinv) {
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int base = sell_base(slice_ptr, i);
  const int len = row_len[i];
  const int4 ti = t_i[i];
  const int vi[4] = {ti.x, ti.y, ti.z, ti.w};
  float w[4], gi[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    w[a] = invm[vi[a]];
#pragma unroll
    for (int r = 0; r < 3; ++r) gi[a][r] = gradC[12 * (size_t)i + 3 * a + r];
  }
  const float al = alpha[i];
  float diag = 0.0f;
  for (int k0 = 0; k0 < len; k0 += 4) {
    const int p = base + 64 * k0;
    const v4i c = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(idx + p));
    const v4u16 dd = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(desc + p));
    const int cj[4] = {c.x, c.y, c.z, c.w};
    const uint32_t dq[4] = {dd.x, dd.y, dd.z, dd.w};
    // gathers of every shared vertex's gradient first, then the arithmetic
    float g[4][4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t d = dq[q];
      const bool all = (d & kDescAll) != 0;
      const int cnt = (d & kDescDiag) ? 0 : all ? 4 : (int)((d >> 12) & 3u);
      const float* gj = gradC + 12 * (size_t)cj[q];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int b = all ? (int)((d >> (2 * e)) & 3u) : (int)((d >> (4 * e + 2)) & 3u);
        if (e < cnt) {
          g[q][e][0] = gj[3 * b];
          g[q][e][1] = gj[3 * b + 1];
          g[q][e][2] = gj[3 * b + 2];
        } else {
          g[q][e][0] = g[q][e][1] = g[q][e][2] = 0.0f;
        }
      }
    }
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t d = dq[q];
      v[q] = 0.0f;
      if (d == kDescDiag) {
        float dv = w[0] * dot3(gi[0], gi[0]);
        dv = fmaf(w[1], dot3(gi[1], gi[1]), dv);
        dv = fmaf(w[2], dot3(gi[2], gi[2]), dv);
        dv = fmaf(w[3], dot3(gi[3], gi[3]), dv);
        v[q] = dv + al;
        diag = v[q];
      } else {
        const bool all = (d & kDescAll) != 0;
        const int cnt = all ? 4 : (int)((d >> 12) & 3u);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (e < cnt) {
            const int a = all ? e : (int)((d >> (4 * e)) & 3u);
            const float ga[3] = {a == 0 ? gi[0][0] : a == 1 ? gi[1][0] : a == 2 ? gi[2][0] : gi[3][0],
                                 a == 0 ? gi[0][1] : a == 1 ? gi[1][1] : a == 2 ? gi[2][1] : gi[3][1],
                                 a == 0 ? gi[0][2] : a == 1 ? gi[1][2] : a == 2 ? gi[2][2] : gi[3][2]};
            const float wa = a == 0 ? w[0] : a == 1 ? w[1] : a == 2 ? w[2] : w[3];
            v[q] = fmaf(wa, dot3(ga, g[q][e]), v[q]);
          }
        }
      }
    }
    *reinterpret_cast<float4*>(val + p) = make_float4(v[0], v[1], v[2], v[3]);
  }
  dinv[i] = fabsf(diag) > 1e-30f ? 1.0f / diag : 0.0f;
}

__global__ void __launch_bounds__(kBlock)
k_scatter(int nv, const int* __restrict__ vptr, const int* __restrict__ vinc,
          const float* __restrict__ invm, const float* __restrict__ gradC,
          const float* __restrict__ dlam, float* __restrict__ dpos) {
  int v = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (v >= nv) return;
  const float w = invm[v];
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
    const int e = vinc[k];
    const int t = e >> 2;
    const float s = w * dlam[t];
    const float* g = gradC + 12 * (size_t)t + 3 * (e & 3);
    a0 = fmaf(s, g[0], a0);
    a1 = fmaf(s, g[1], a1);
    a2 = fmaf(s, g[2], a2);
  }
  dpos[3 * (size_t)v + 0] = a0;
  dpos[3 * (size_t)v + 1] = a1;
  dpos[3 * (size_t)v + 2] = a2;
}

__global__ void __launch_bounds__(kBlock)
k_add_inplace(int n, const float* __restrict__ d, float* __restrict__ y) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = y[i] + d[i];
}

__global__ void __launch_bounds__(kBlock)
k_apply_dpos(int nv, float* __restrict__ x, const float* __restrict__ dpos,
             const float* __restrict__ invm, float omega) {
  int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= nv || !(invm[v] > 0.0f)) return;
#pragma unroll
  for (int r = 0; r < 3; ++r) x[3 * (size_t)v + r] = fmaf(omega, dpos[3 * (size_t)v + r], x[3 * (size_t)v + r]);
}

// f_ext = 0 except the dragged vertex: k (target - x)
__global__ void __launch_bounds__(kBlock)
k_spring_drag(int nv, const float* __restrict__ x, float* __restrict__ f, int vertex, float tx,
              float ty, float tz, float k) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= 3 * nv) return;
  const int v = e / 3, c = e % 3;
  const float t = c == 0 ? tx : (c == 1 ? ty : tz);
  f[e] = v == vertex ? k * (t - x[e]) : 0.0f;
}

__global__ void __launch_bounds__(kBlock)
k_predict(int nv, float* __restrict__ x, float* __restrict__ vel, float* __restrict__ x_old,
          const float* __restrict__ invm, const float* __restrict__ fext, float gx, float gy,
          float gz, float dt, float damp, int hold_pinned) {
  int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= nv) return;
  const float g[3] = {gx, gy, gz};
  const bool pinned = hold_pinned && !(invm[v] > 0.0f);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    size_t q = 3 * (size_t)v + r;
    float xv = x[q];
    if (pinned) {
      vel[q] = 0.0f;
      x_old[q] = xv;
      continue;
    }
    float vv = vel[q] + dt * g[r];
    if (fext) vv = vv + dt * invm[v] * fext[q];
    vv = vv * damp;
    vel[q] = vv;
    x_old[q] = xv;
    x[q] = xv + dt * vv;
  }
}

__global__ void __launch_bounds__(kBlock)
k_update_vel(int nv, const float* __restrict__ x, const float* __restrict__ x_old,
             float* __restrict__ vel, float dt, float damp) {
  int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= 3 * nv) return;
  vel[q] = damp * (x[q] - x_old[q]) / dt;
}

// ------------------------------------------------------------------------------------------
// SELL-64x4 sparse kernels. One thread per row; the wavefront owns one slice, every lane
// loads four columns and four values (16 bytes each) at a time, so a wavefront's load is
// one contiguous 1 KiB segment; the matrix streams are read with non-temporal loads (each
// byte is used once per pass) to keep x resident in L2. Rows are summed sequentially in CSR
// order (the order of the reference's row-per-thread kernels).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float sell_row(int i, const int* __restrict__ slice_ptr,
                                          const int* __restrict__ idx,
                                          const float* __restrict__ val,
                                          const float* __restrict__ x) {
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  float acc = 0.0f;
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    const v4i c = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(idx + p));
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(val + p));
    acc = fmaf(v.x, x[c.x], acc);
    acc = fmaf(v.y, x[c.y], acc);
    acc = fmaf(v.z, x[c.z], acc);
    acc = fmaf(v.w, x[c.w], acc);
  }
  return acc;
}

// The same row product with FP64 products and sums (the FP32 x2 FP32 products are exact in
// FP64), x in FP32 or FP64.
template <typename XT>
__device__ __forceinline__ double sell_row_d(int i, const int* __restrict__ slice_ptr,
                                             const int* __restrict__ idx,
                                             const float* __restrict__ val,
                                             const XT* __restrict__ x) {
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  double acc = 0.0;
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    const v4i c = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(idx + p));
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(val + p));
    acc = fma((double)v.x, (double)x[c.x], acc);
    acc = fma((double)v.y, (double)x[c.y], acc);
    acc = fma((double)v.z, (double)x[c.z], acc);
    acc = fma((double)v.w, (double)x[c.w], acc);
  }
  return acc;
}

// MODE 0: y = A x;  MODE 1: y = x + (omega dinv) (b - A x);  MODE 2: y = b - A x;
// MODE 3 (Chebyshev step): y = x + c1 (x - xp) + (c2 dinv) (b - A x), (c1, c2) = cc[0..1],
// xp = the previous iterate (NULL: zero)
template <int MODE>
__global__ void __launch_bounds__(kBlock)
k_sell(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
       const float* __restrict__ val, const float* __restrict__ x, float* y,
       const float* __restrict__ dinv, const float* __restrict__ b, float omega, const float* __restrict__ wp,
       const int* __restrict__ done, const float* xp = nullptr,
       const float* __restrict__ cc = nullptr) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const float ax = sell_row(i, slice_ptr, idx, val, x);
  if (i >= n) return;
  if (MODE == 0) y[i] = ax;
  else if (MODE == 1) y[i] = fmaf(omega * dinv[i], b[i] - ax, x[i]);
  else if (MODE == 2) y[i] = b[i] - ax;
  else {
    const float xi = x[i];
    const float u = fmaf(cc[0], xi - (xp ? xp[i] : 0.0f), xi);
    y[i] = fmaf(cc[1] * dinv[i], b[i] - ax, u);
  }
}

// Power iteration on D^-1 A (symmetric in the D inner product): w = D^-1 A v and the
// partials v.Av (slot 0), v.Dv (slot 1), w.Dw = (Av)^2 / d (slot 2).
__global__ void __launch_bounds__(kBlock)
k_power_step(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
             const float* __restrict__ val, const float* __restrict__ dinv,
             const float* __restrict__ v, float* __restrict__ w, RedWS ws) {
  double vav = 0.0, vdv = 0.0, wdw = 0.0;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = sell_row(i, slice_ptr, idx, val, v);
    if (i < n) {
      const float di = dinv[i];
      w[i] = di * ax;
      vav = (double)v[i] * ax;
      if (di > 0.0f) {
        vdv = (double)v[i] * v[i] / di;
        wdw = (double)ax * ax * di;
      }
    }
  }
  const double s0 = block_sum(vav);
  const double s1 = block_sum(vdv);
  const double s2 = block_sum(wdw);
  store_partial(ws, 0, s0);
  store_partial(ws, 1, s1);
  store_partial(ws, 2, s2);
}

// pw[0] = lambda = v.Av / v.Dv, pw[1] = ||w||_D, pw[2] = the previous iteration's lambda
__global__ void __launch_bounds__(kBlock)
k_power_fin(RedWS ws, int nparts, double* __restrict__ pw) {
  double t[3];
  sum_partials_k<3>(ws, 0, nparts, t);
  const double vav = t[0], vdv = t[1], wdw = t[2];
  if (threadIdx.x != 0) return;
  pw[2] = pw[0];
  pw[0] = vdv > 0.0 ? vav / vdv : 0.0;
  pw[1] = sqrt(wdw);
}

// k_power_fin and k_power_norm in one launch for levels with few partials: every block sums
// the same partials in the same order (identical totals), block 0 stores the estimates.
__global__ void __launch_bounds__(kBlock)
k_power_fin_norm(int n, RedWS ws, int nparts, const float* __restrict__ w,
                 float* __restrict__ v, double* __restrict__ pw) {
  double t[3];
  sum_partials_k<3>(ws, 0, nparts, t);
  const double vav = t[0], vdv = t[1], wdw = t[2];
  __shared__ double nrm_s;
  if (threadIdx.x == 0) {
    nrm_s = sqrt(wdw);
    if (blockIdx.x == 0) {
      pw[2] = pw[0];
      pw[0] = vdv > 0.0 ? vav / vdv : 0.0;
      pw[1] = nrm_s;
    }
  }
  __syncthreads();
  const double nrm = nrm_s;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !(nrm > 0.0) || !(nrm < 1e300)) return;
  v[i] = (float)(w[i] / nrm);
}

// v = w / ||w||_D (v unchanged when the norm is 0 or not finite)
__global__ void __launch_bounds__(kBlock)
k_power_norm(int n, const float* __restrict__ w, float* __restrict__ v,
             const double* __restrict__ pw) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const double nrm = pw[1];
  if (i >= n || !(nrm > 0.0) || !(nrm < 1e300)) return;
  v[i] = (float)(w[i] / nrm);
}

// deterministic start vector with entries of both signs
__global__ void __launch_bounds__(kBlock)
k_power_init(int n, float* __restrict__ v) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t h = (uint32_t)i * 2654435761u;
  h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
  v[i] = 0.5f + (float)(h & 0xFFFFu) * (1.0f / 65536.0f) - ((h >> 16) & 1u ? 1.0f : 0.0f);
}

// Level weights from the power-iteration estimate (raised by a quarter while the iteration is
// still moving: the last two Rayleigh quotients differ by more than 2 %).
//  Jacobi:    omega_l = target / (safety lambda).
//  Chebyshev: on [a, b] = [safety lambda / ratio, safety lambda], theta = (b + a) / 2,
//             delta = (b - a) / 2, sigma = theta / delta: the first step is Jacobi with
//             omega_l = 1 / theta, step k >= 2 uses c1 = rho_k rho_{k-1}, c2 = 2 rho_k / delta with
//             rho_1 = 1 / sigma, rho_k = 1 / (2 sigma - rho_{k-1}).
// Not usable lambda: the fixed weight (and Chebyshev steps that reduce to it, c1 = 0).
// Chebyshev coefficient pairs kept per level: steps 2 .. kChebSteps + 1 of a smoothing phase
// (a phase of at most kChebSteps + 1 steps; the per-level step count is capped at kChebSteps)
constexpr int kChebSteps = 16;

__global__ void k_power_omega(const double* __restrict__ pw, float target, float safety,
                              float fallback, int cheb, float ratio, float* __restrict__ omega_l,
                              float* __restrict__ cheb_l) {
  if (threadIdx.x != 0) return;
  double lam = pw[0];
  if (!(fabs(lam - pw[2]) <= 0.02 * lam)) lam *= 1.25;
  const bool ok = lam > 1e-6 && lam < 1e30;
  if (!cheb) {
    omega_l[0] = ok ? (float)(target / (safety * lam)) : fallback;
    return;
  }
  if (!ok) {
    omega_l[0] = fallback;
    for (int k = 0; k < kChebSteps; ++k) { cheb_l[2 * k] = 0.0f; cheb_l[2 * k + 1] = fallback; }
    return;
  }
  const double b = safety * lam, a = b / ratio;
  const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
  omega_l[0] = (float)(1.0 / theta);
  double rho = 1.0 / sigma;
  for (int k = 0; k < kChebSteps; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    cheb_l[2 * k] = (float)(rn * rho);
    cheb_l[2 * k + 1] = (float)(2.0 * rn / delta);
    rho = rn;
  }
}

__global__ void k_set_float(float* __restrict__ x, float v) {
  if (threadIdx.x == 0) x[0] = v;
}

// Several Jacobi sweeps on a diagonal-only matrix in one pass (rows are independent, so this
// is the same arithmetic as `sweeps` separate launches).
__global__ void __launch_bounds__(kBlock)
k_jacobi_diag(int n, const int* __restrict__ slice_ptr, const float* __restrict__ val,
              const float* __restrict__ dinv, const float* __restrict__ b,
              float* __restrict__ x, float omega, const float* __restrict__ wp, int sweeps, const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float a = val[sell_base(slice_ptr, i)];
  const float wd = omega * dinv[i];
  const float bi = b[i];
  float xi = x[i];
  for (int s = 0; s < sweeps; ++s) xi = fmaf(wd, bi - fmaf(a, xi, 0.0f), xi);
  x[i] = xi;
}

// All `sweeps` Jacobi sweeps of a small level (n <= kSmallRows) in one workgroup: x lives in
// LDS (two buffers), a __syncthreads separates sweeps. Each row does the same arithmetic as
// k_sell<1>, so the result equals `sweeps` separate launches bit for bit.
constexpr int kSmallRows = 4096;
constexpr int kSmallBlock = 1024;

__global__ void __launch_bounds__(kSmallBlock)
k_jacobi_small(int n, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
               const float* __restrict__ val, const float* __restrict__ dinv,
               const float* __restrict__ b, float* __restrict__ x, float omega, const float* __restrict__ wp, int sweeps,
               const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  __shared__ float xs[2][kSmallRows];
  const int n_slices = (n + 63) >> 6;
  for (int i = threadIdx.x; i < 64 * n_slices; i += kSmallBlock) {
    xs[0][i] = i < n ? x[i] : 0.0f;
    xs[1][i] = 0.0f;
  }
  __syncthreads();
  int cur = 0;
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int i = threadIdx.x; i < 64 * n_slices; i += kSmallBlock) {
      const int p0 = sell_base(slice_ptr, i);
      const int w = sell_width(slice_ptr, i);
      float acc = 0.0f;
      for (int k = 0; k < w; ++k) {
        const int p = p0 + 64 * (k & ~3) + (k & 3);
        acc = fmaf(val[p], xs[cur][idx[p]], acc);
      }
      if (i < n) xs[cur ^ 1][i] = fmaf(omega * dinv[i], b[i] - acc, xs[cur][i]);
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int i = threadIdx.x; i < n; i += kSmallBlock) x[i] = xs[cur][i];
}

// Coarsest-level direct solve (FP64): the level's SELL values are expanded into a dense
// column-major n x n matrix, factored and inverted by rocSOLVER (dpotrf + dpotri, lower
// triangle), mirrored to a full symmetric inverse, and every V-cycle applies it as a dense
// mat-vec with FP64 accumulation.
__global__ void __launch_bounds__(kBlock)
k_dense_zero(int64_t len, double* __restrict__ D) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < len) D[i] = 0.0;
}

// One wavefront per row after k_dense_zero: lane k stores entry k of row i (a row's nonzero
// columns are distinct; zero padding, whose column is the row itself, is skipped).
__global__ void __launch_bounds__(kBlock)
k_dense_from_sell(int n, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
                  const float* __restrict__ val, double* __restrict__ D) {
  const int i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (i >= n) return;
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  for (int k = threadIdx.x & 63; k < w; k += 64) {
    const int p = p0 + 64 * (k & ~3) + (k & 3);
    const float v = val[p];
    if (v != 0.0f) D[(size_t)idx[p] * n + i] = (double)v;   // column idx, row i
  }
}

// Newton-Schulz acceptance on the device: X replaces the inverse when ||I - A X||_F^2 < 0.01
// (f2 from k_fin_sum); flag = 1 accepted, 2 rejected (the previous inverse stays).
__global__ void __launch_bounds__(kBlock)
k_ns_accept(int64_t len, const double* __restrict__ f2, const double* __restrict__ X,
            double* __restrict__ D, int* __restrict__ flag) {
  const bool ok = *f2 < 0.01;   // false for a diverged (non-finite) iteration too
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) *flag = ok ? 1 : 2;
  if (ok && i < len) D[i] = X[i];
}

// R = I (column-major n x n)
__global__ void __launch_bounds__(kBlock)
k_set_identity(int n, double* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  R[t] = (t % n == t / n) ? 1.0 : 0.0;
}

// partial sums of squares of len doubles (slot 0)
__global__ void __launch_bounds__(kBlock)
k_sumsq(int64_t len, const double* __restrict__ x, RedWS ws) {
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < len;
       t += (int64_t)gridDim.x * kBlock)
    acc += x[t] * x[t];
  store_partial(ws, 0, block_sum(acc));
}

// C = alpha A B + beta C for n x n column-major FP64 matrices (beta = 0: C is not read) on
// the FP64 matrix cores. A block owns one 16 x 16 tile of C; its four wavefronts split the k
// range in four consecutive parts and each runs v_mfma_f64_16x16x4 along its part: in the
// m-th MFMA of a 16-wide k step, lane l supplies A(i0 + l % 16, k) and B(k, j0 + l % 16) with
// k = k0 + 4 (l / 16) + m, and accumulator r of lane l is C(i0 + l / 16 + 4 r, j0 + l % 16)
// (the layout measured on gfx950 with a one-hot B operand). The next step's operands are
// loaded before the current step's MFMAs issue; the split gives each SIMD several wavefronts
// to cover the L2 latency at the coarsest level's sizes. The four partial tiles are added in
// wavefront order through LDS, so the result is deterministic. Grid (ceil(n/16), ceil(n/16));
// rows, columns and k past n read zeros.
constexpr int kGemmTile = 16;
constexpr int kGemmSplit = kBlock / 64;
typedef double v4f64 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void dgemm_operands(int n, const double* __restrict__ A,
                                               const double* __restrict__ Bc, int ia, bool ra,
                                               bool rb, int kb, double (&a)[4],
                                               double (&b)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {   // loads from clamped (in-bounds) addresses, then selects:
    const int k = kb + m;         // no branches around the loads
    const bool in = k < n;
    const int kc = in ? k : n - 1;
    const double av = A[(size_t)kc * n + (ra ? ia : 0)];
    const double bv = Bc[kc];
    a[m] = (ra && in) ? av : 0.0;
    b[m] = (rb && in) ? bv : 0.0;
  }
}

__global__ void __launch_bounds__(kBlock)
k_dgemm(int n, double alpha, const double* __restrict__ A, const double* __restrict__ B,
        double beta, double* __restrict__ C) {
  __shared__ v4f64 part[kGemmSplit - 1][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i0 = blockIdx.x * kGemmTile, j0 = blockIdx.y * kGemmTile;
  const int li = lane & 15, g = lane >> 4;
  const int ia = i0 + li, jb = j0 + li;
  const bool ra = ia < n, rb = jb < n;
  const double* Bc = B + (size_t)(rb ? jb : 0) * n;
  const int steps = (n + 15) / 16;
  const int per = (steps + kGemmSplit - 1) / kGemmSplit;
  const int s0 = w * per, s1 = min(steps, s0 + per);
  v4f64 acc = {0.0, 0.0, 0.0, 0.0};
  if (s0 < s1) {   // uniform across the wavefront
    double a[4], b[4], an[4], bn[4];
    dgemm_operands(n, A, Bc, ia, ra, rb, 16 * s0 + 4 * g, a, b);
    for (int st = s0; st < s1; ++st) {
      dgemm_operands(n, A, Bc, ia, ra, rb, 16 * (st + 1) + 4 * g, an, bn);
#pragma unroll
      for (int m = 0; m < 4; ++m)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[m], acc, 0, 0, 0);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        a[m] = an[m];
        b[m] = bn[m];
      }
    }
  }
  if (w > 0) part[w - 1][lane] = acc;
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int q = 0; q < kGemmSplit - 1; ++q) acc += part[q][lane];
  const int j = j0 + li;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + g + 4 * r;
    if (i < n && j < n) {
      double* c = C + (size_t)j * n + i;
      *c = beta == 0.0 ? alpha * acc[r] : fma(alpha, acc[r], beta * *c);
    }
  }
}

// out = max_i sum_j |A_ij| (one block; n x n column-major)
__global__ void __launch_bounds__(kBlock)
k_inf_norm(int n, const double* __restrict__ A, double* __restrict__ out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) {
    double r = 0.0;
    for (int j = 0; j < n; ++j) r += fabs(A[(size_t)j * n + i]);
    m = fmax(m, r);
  }
  __shared__ double sm[kBlock];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// X = I / *nrm (column-major n x n)
__global__ void __launch_bounds__(kBlock)
k_set_scaled_identity(int n, const double* __restrict__ nrm, double* __restrict__ X) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  X[t] = (t % n == t / n) ? 1.0 / *nrm : 0.0;
}

// x = D b for a dense symmetric D (column-major): one wavefront per row i reads column i
// (= row i) contiguously, FP64 accumulation, fixed shuffle-tree reduction.
__global__ void __launch_bounds__(kBlock)
k_dense_matvec(int n, const double* __restrict__ D, const float* __restrict__ b,
               float* __restrict__ x, const int* __restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* col = D + (size_t)i * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += col[j] * (double)b[j];
  acc = wave_sum(acc);
  if (lane == 0) x[i] = (float)acc;
}

__global__ void __launch_bounds__(kBlock)
k_diag_inv(int n, const int* __restrict__ diag_pos, const float* __restrict__ val,
           float* __restrict__ dinv) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int p = diag_pos[i];
  float d = p >= 0 ? val[p] : 0.0f;
  dinv[i] = fabsf(d) > 1e-30f ? 1.0f / d : 0.0f;
}

__global__ void __launch_bounds__(kBlock)
k_refill(int nc, const int* __restrict__ off, const int* __restrict__ fpos,
         const int* __restrict__ cpos, const float* __restrict__ fine, float* __restrict__ coarse) {
  int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (c >= nc) return;
  // contributions in batches of 8: all index and value loads of a batch are issued before
  // the adds, which stay in the map's order (the sum is the same as one add per load)
  float v = 0.0f;
  int k = off[c];
  const int k1 = off[c + 1];
  for (; k + 8 <= k1; k += 8) {
    int q[8];
    float f[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = fpos[k + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) f[u] = fine[q[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) v = v + f[u];
  }
  for (; k < k1; ++k) v = v + fine[fpos[k]];
  coarse[cpos[c]] = v;
}

// Sum over the members of aggregate I (ascending) of f(i) = r_i (or sd_i r_i) in the member
// order: the member indices of a batch of up to kAggBatch are loaded first, then their values,
// then summed, so an aggregate (at most 8 members, a few more after singleton merges) costs two
// dependent load rounds instead of two per member. Members past the end re-read the last one
// and are dropped, so the sum equals the one-member-at-a-time loop bit for bit.
constexpr int kAggBatch = 8;
template <bool SCALED>
__device__ __forceinline__ float agg_sum(int I, const int* __restrict__ mptr,
                                         const int* __restrict__ mem,
                                         const float* __restrict__ sd,
                                         const float* __restrict__ r) {
  const int m0 = mptr[I], m1 = mptr[I + 1];
  float s = 0.0f;
  for (int mb = m0; mb < m1; mb += kAggBatch) {
    int id[kAggBatch];
#pragma unroll
    for (int j = 0; j < kAggBatch; ++j) id[j] = mem[min(mb + j, m1 - 1)];
    float rv[kAggBatch], sv[kAggBatch];
#pragma unroll
    for (int j = 0; j < kAggBatch; ++j) {
      rv[j] = r[id[j]];
      sv[j] = SCALED ? sd[id[j]] : 1.0f;
    }
#pragma unroll
    for (int j = 0; j < kAggBatch; ++j) {
      if (mb + j < m1) s = SCALED ? fmaf(sv[j], rv[j], s) : s + rv[j];
    }
  }
  return s;
}

// Prolongation in push form, fused into the dense coarsest product: coarse row I adds its
// correction to the members of its aggregate on the next finer level, x_j = fmaf((sd_j isd_I)
// g, xc_I, x_j) for j in mem[mptr[I] .. mptr[I + 1]) -- the arithmetic of k_prolong_h on the
// same values, so the result is bit-identical to a separate prolongation launch. Every fine row
// belongs to exactly one aggregate, so no two lanes write one x_j. (Pushing from the sweep
// kernels of the other coarse levels was measured slower than their row-parallel k_prolong_h:
// each sweep thread then waits on scattered read-modify-writes, 3.5-6 us per launch against
// 3.3-5 us for the separate pass.) x == nullptr: nothing to push.
struct Push {
  const int* mptr;
  const int* mem;
  const float* sd_f;
  const float* isd_c;
  float* x;
  float g;
};

// b_c[I] = sum of r over the members of aggregate I (ascending); with dinv_c given, also the
// coarse level's first Jacobi sweep from zero, x_c = (omega dinv_c) b_c.
__global__ void __launch_bounds__(kBlock)
k_restrict(int nc, const int* __restrict__ mptr, const int* __restrict__ mem,
           const float* __restrict__ r, float* __restrict__ bc, const float* __restrict__ dinv_c,
           float* __restrict__ xc, float omega, const float* __restrict__ wp, const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  int I = blockIdx.x * kBlock + threadIdx.x;
  if (I >= nc) return;
  const float s = agg_sum<false>(I, mptr, mem, nullptr, r);
  bc[I] = s;
  if (xc) xc[I] = (omega * dinv_c[I]) * s;
}

__global__ void __launch_bounds__(kBlock)
k_prolong(int n, const int* __restrict__ agg, const float* __restrict__ xc,
          float* __restrict__ x, float g, const int* __restrict__ done) {
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = fmaf(g, xc[agg[i]], x[i]);   // g = 1: x + xc, rounded once
}

// Second coarse cycle of a W-cycle: y = x (save), e = e + s (combine), with the done flag.
__global__ void __launch_bounds__(kBlock)
k_copy_d(int n, const float* __restrict__ x, float* __restrict__ y, const int* __restrict__ done) {
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = x[i];
}
__global__ void __launch_bounds__(kBlock)
k_add_d(int n, const float* __restrict__ s, float* __restrict__ e, const int* __restrict__ done) {
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) e[i] = e[i] + s[i];
}
// x = omega b: the first sweep from zero on a unit-diagonal (scaled) level
__global__ void __launch_bounds__(kBlock)
k_scale_d(int n, float omega, const float* __restrict__ wp, const float* __restrict__ b, float* __restrict__ x,
          const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = omega * b[i];
}

// x = (omega dinv) b : the first Jacobi sweep from x = 0
__global__ void __launch_bounds__(kBlock)
k_first_sweep(int n, const float* __restrict__ dinv, const float* __restrict__ b,
              float* __restrict__ x, float omega, const float* __restrict__ wp, const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = (omega * dinv[i]) * b[i];
}

// ------------------------------------------------------------------------------------------
// BLAS-1
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_blas1(int op, int n, float alpha, const float* __restrict__ a, float* __restrict__ y) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case 0: y[i] = fmaf(alpha, a[i], y[i]); break;      // axpy  y += alpha x
    case 1: y[i] = fmaf(alpha, y[i], a[i]); break;      // xpay  y = x + alpha y
    case 2: y[i] = a[i]; break;                         // copy
    case 3: y[i] = alpha; break;                        // fill
    case 4: y[i] = y[i] * alpha; break;                 // scale
    case 5: y[i] = a[i] - y[i]; break;                  // negate_and_add  r = b - r
  }
}

__global__ void __launch_bounds__(kBlock)
k_dot(int n, const float* __restrict__ a, const float* __restrict__ b, RedWS ws) {
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    acc += (double)a[i] * (double)b[i];
  double bs = block_sum(acc);
  store_partial(ws, 0, bs);
}

// ------------------------------------------------------------------------------------------
// PCG (cons/mgpcg_gpu.py:261-329) with the scalars kept on the device.
// ------------------------------------------------------------------------------------------
struct PcgState {
  double rz, pAp, alpha, beta, rr, bb, tol2, tol;
  double rr0, rr_guess, rr_alt;   // initial residuals (started from a guess)
  double rr_true;                 // FP64 ||b - A x||^2 of the returned x (verification)
  int done;     // 0 running, 1 converged, 2 |pAp| < 1e-30, 3 |rz| < 1e-30
  int iters;    // completed iterations (the reference's cg_iters)
  int k;        // iteration index
  int relative;
  int start;    // 0: x = 0, 1: the guess in x, 2: the alternative guess
  int verified; // 1: rr_true was computed on the device after convergence (graph tail)
};

__global__ void k_pcg_reset(PcgState* st, double tol, int relative) {
  st->done = 0; st->iters = 0; st->k = 0; st->relative = relative; st->tol = tol;
  st->rz = st->pAp = st->alpha = st->beta = st->rr = st->bb = 0.0;
  st->rr0 = st->rr_guess = st->rr_alt = 0.0;
  st->rr_true = -1.0;
  st->start = 0;
  st->verified = 0;
  st->tol2 = tol * tol;
}

// Started from a guess: partials of b.b (slot 0), r.r (slot 1) and q.q (slot 2), where
// r = b - A x_guess and q = b - A x_alt were written by the residual pass; with `energy`, also
// x.(b + r) (slot 3) and x_alt.(b + q) (slot 4), from which phi(x) = x.A x / 2 - b.x =
// -x.(b + r) / 2, the CG energy (phi(x) - phi(x*) = ||x - x*||_A^2 / 2).
__global__ void __launch_bounds__(kBlock)
k_pcg_guess_norms(int n, const float* __restrict__ b, const float* __restrict__ r,
                  const float* __restrict__ q, const float* __restrict__ x,
                  const float* __restrict__ x_alt, int energy, RedWS ws) {
  double bb = 0.0, rr = 0.0, qq = 0.0, ea = 0.0, eb = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double bi = b[i], ri = r[i];
    bb += bi * bi;
    rr += ri * ri;
    if (energy) ea += (double)x[i] * (bi + ri);
    if (q) {
      const double qi = q[i];
      qq += qi * qi;
      if (energy) eb += (double)x_alt[i] * (bi + qi);
    }
  }
  const double b0 = block_sum(bb);
  const double b1 = block_sum(rr);
  const double b2 = block_sum(qq);
  store_partial(ws, 0, b0);
  store_partial(ws, 1, b1);
  store_partial(ws, 2, b2);
  if (energy) {
    const double b3 = block_sum(ea);
    const double b4 = block_sum(eb);
    store_partial(ws, 3, b3);
    store_partial(ws, 4, b4);
  }
}

// Picks the start among the guess, the alternative guess and zero: by the smallest residual,
// or (energy) by the lowest CG energy phi, i.e. the smallest A-norm error (phi(0) = 0). The
// guess wins a tie with the alternative; zero wins a tie and any non-finite value. The stopping
// test stays relative to ||b|| (or absolute), as from a zero start.
__global__ void __launch_bounds__(kBlock)
k_pcg_fin_guess(RedWS ws, int nparts, int has_alt, int energy, PcgState* st) {
  double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (energy) {
    sum_partials_k<5>(ws, 0, nparts, t);
  } else {
    double t3[3];
    sum_partials_k<3>(ws, 0, nparts, t3);
    t[0] = t3[0]; t[1] = t3[1]; t[2] = t3[2];
  }
  const double bb = t[0], ra = t[1], rb = t[2];
  if (threadIdx.x != 0) return;
  st->bb = bb;
  if (st->relative) st->tol2 = st->tol * st->tol * bb;
  st->rr_guess = ra;
  st->rr_alt = has_alt ? rb : 0.0;
  int start = 1;
  if (energy) {
    const double pa = -0.5 * t[3], pb = -0.5 * t[4];
    double best = pa;
    if (has_alt && pb < best) { start = 2; best = pb; }
    if (!(best < 0.0) || !(ra < 1e300) || (start == 2 && !(rb < 1e300))) start = 0;
  } else {
    double best = ra;
    if (has_alt && rb < best) { start = 2; best = rb; }
    if (!(best < bb)) start = 0;   // also catches a non-finite guess
  }
  st->start = start;
  st->rr0 = start == 0 ? bb : start == 2 ? rb : ra;
}

// Applies the chosen start: x, r (and the fused first sweep xv = (omega dinv0) r).
__global__ void __launch_bounds__(kBlock)
k_pcg_select(int n, const float* __restrict__ b, float* __restrict__ x, float* __restrict__ r,
             const float* __restrict__ x_alt, const float* __restrict__ q,
             const float* __restrict__ dinv0, float* __restrict__ xv, float omega, const float* __restrict__ wp,
             const PcgState* __restrict__ st) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int start = st->start;
  float ri;
  if (start == 0) { x[i] = 0.0f; ri = b[i]; r[i] = ri; }
  else if (start == 2) { x[i] = x_alt[i]; ri = q[i]; r[i] = ri; }
  else ri = r[i];
  if (xv) xv[i] = (omega * dinv0[i]) * ri;
}

// FP64 residual of the FP32 iterate: products and sums in double over the FP32 A, x, b;
// partial ||b - A x||^2 in slot 0
__global__ void __launch_bounds__(kBlock)
k_true_resid(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
             const float* __restrict__ val, const float* __restrict__ x,
             const float* __restrict__ b, RedWS ws) {
  double acc = 0.0;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const int p0 = sell_base(slice_ptr, i);
    const int w = sell_width(slice_ptr, i);
    double ax = 0.0;
    for (int k0 = 0; k0 < w; k0 += 4) {
      const int p = p0 + 64 * k0;
      const v4i c = *reinterpret_cast<const v4i*>(idx + p);
      const v4f v = *reinterpret_cast<const v4f*>(val + p);
      ax = fma((double)v.x, (double)x[c.x], ax);
      ax = fma((double)v.y, (double)x[c.y], ax);
      ax = fma((double)v.z, (double)x[c.z], ax);
      ax = fma((double)v.w, (double)x[c.w], ax);
    }
    if (i < n) {
      const double ri = (double)b[i] - ax;
      acc = ri * ri;
    }
  }
  store_partial(ws, 0, block_sum(acc));
}

__global__ void __launch_bounds__(kBlock)
k_fin_true(RedWS ws, int nparts, PcgState* st) {
  const double t = sum_partials(ws, 0, nparts);
  if (threadIdx.x == 0) st->rr_true = t;
}

// The verification tail of a captured PCG replay: runs only once the solve has converged
// (done == 1, a value every thread reads alike), then marks the state verified.
__global__ void __launch_bounds__(kBlock)
k_fin_true_gated(RedWS ws, int nparts, PcgState* st) {
  if (st->done != 1 || st->verified) return;
  const double t = sum_partials(ws, 0, nparts);
  if (threadIdx.x == 0) {
    st->rr_true = t;
    st->verified = 1;
  }
}

__global__ void __launch_bounds__(kBlock)
k_x_from64_gated(int n, const double* __restrict__ x64, float* __restrict__ x,
                 const PcgState* __restrict__ st) {
  if (st->done != 1 || st->verified) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = (float)x64[i];
}

// Restart residual of the FP64 iterate: r = b - A x64 (FP64 products and sums), rounded once
__global__ void __launch_bounds__(kBlock)
k_resid64(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
          const float* __restrict__ val, const double* __restrict__ x64,
          const float* __restrict__ b, float* __restrict__ r) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const double ax = sell_row_d(i, slice_ptr, idx, val, x64);
  if (i < n) r[i] = (float)((double)b[i] - ax);
}

// FP64 iterate from the start the device selected: the FP32 start on the first round, and on
// a restart only when the zero start was taken (otherwise x64 already holds the iterate)
__global__ void __launch_bounds__(kBlock)
k_x64_from_start(int n, const float* __restrict__ x, double* __restrict__ x64, int restart,
                 const PcgState* __restrict__ st) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || (restart && st->start != 0)) return;
  x64[i] = (double)x[i];
}

// the FP32 result of the FP64 iterate
__global__ void __launch_bounds__(kBlock)
k_x_from64(int n, const double* __restrict__ x64, float* __restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = (float)x64[i];
}

// r = b, x = 0, xv = (omega dinv0) r (first pre-smooth sweep, when wanted), partial b.b
__global__ void __launch_bounds__(kBlock)
k_pcg_init(int n, const float* __restrict__ b, float* __restrict__ x, float* __restrict__ r,
           const float* __restrict__ dinv0, float* __restrict__ xv, float omega, const float* __restrict__ wp, RedWS ws,
           double* __restrict__ x64) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float bi = b[i];
    r[i] = bi;
    x[i] = 0.0f;
    if (x64) x64[i] = 0.0;
    if (xv) xv[i] = (omega * dinv0[i]) * bi;
    acc += (double)bi * (double)bi;
  }
  store_partial(ws, 0, block_sum(acc));
}

// partial r.z
__global__ void __launch_bounds__(kBlock)
k_pcg_rz(int n, const float* __restrict__ r, const float* __restrict__ z, RedWS ws,
         const PcgState* st) {
  if (st->done) return;
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    acc += (double)r[i] * (double)z[i];
  store_partial(ws, 0, block_sum(acc));
}

// Ap = A p and partial p.Ap (one row per thread, one slice per wavefront)
// ACC64: products and sums in FP64, Ap rounded once to FP32
template <bool ACC64>
__global__ void __launch_bounds__(kBlock)
k_pcg_spmv_dot(int n, int n_slices, const int* __restrict__ slice_ptr,
               const int* __restrict__ idx, const float* __restrict__ val,
               const float* __restrict__ p, float* __restrict__ Ap, RedWS ws,
               const PcgState* st) {
  if (st->done) return;
  double acc = 0.0;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = ACC64 ? (float)sell_row_d(i, slice_ptr, idx, val, p)
                           : sell_row(i, slice_ptr, idx, val, p);
    if (i < n) {
      Ap[i] = ax;
      acc = (double)p[i] * (double)ax;
    }
  }
  store_partial(ws, 0, block_sum(acc));
}

// x += alpha p, r -= alpha Ap, partial r.r; also xv = (omega dinv0) r for the next V-cycle.
// x64 != NULL: the iterate is the FP64 x64 (x64 += alpha p in FP64; x is left alone)
__global__ void __launch_bounds__(kBlock)
k_pcg_update(int n, float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
             const float* __restrict__ Ap, const float* __restrict__ dinv0,
             float* __restrict__ xv, float omega, const float* __restrict__ wp, RedWS ws,
             const PcgState* st, double* __restrict__ x64, float* __restrict__ bh) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (st->done) return;
  const float al = (float)st->alpha;
  const double ald = st->alpha;
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (x64) x64[i] = fma(ald, (double)p[i], x64[i]);
    else x[i] = fmaf(al, p[i], x[i]);
    float ri = fmaf(-al, Ap[i], r[i]);
    r[i] = ri;
    if (bh) {   // scaled level 0 (dinv0 = isd): k_half_first's b^ and first sweep
      const float v = dinv0[i] * ri;
      bh[i] = v;
      xv[i] = omega * v;
    } else if (xv) {
      xv[i] = (omega * dinv0[i]) * ri;
    }
    acc += (double)ri * (double)ri;
  }
  store_partial(ws, 0, block_sum(acc));
}

// Scalar epilogues of the PCG (cons/mgpcg_gpu.py:286-327), one block.
enum PcgFin { FIN_BB = 0, FIN_RZ0 = 1, FIN_ALPHA = 2, FIN_RR = 3, FIN_BETA = 4 };

template <int KIND>
__global__ void __launch_bounds__(kBlock)
k_pcg_fin(RedWS ws, int nparts, PcgState* st) {
  if (KIND != FIN_BB && st->done) return;
  double tot = sum_partials(ws, 0, nparts);
  if (threadIdx.x != 0) return;
  if (KIND == FIN_BB) {
    st->bb = tot;
    if (st->relative) st->tol2 = st->tol * st->tol * tot;
  } else if (KIND == FIN_RZ0) {
    st->rz = tot;
  } else if (KIND == FIN_ALPHA) {
    st->pAp = tot;
    if (fabs(tot) < 1e-30) st->done = 2;
    else st->alpha = st->rz / tot;
  } else if (KIND == FIN_RR) {
    st->rr = tot;
    if (tot < st->tol2) {
      st->done = 1;
      st->iters = st->k + 1;
    }
  } else {   // FIN_BETA: rz_new = tot; breakdown test on the previous rz
    st->iters = st->k + 1;
    if (fabs(st->rz) < 1e-30) {
      st->done = 3;
    } else {
      st->beta = tot / st->rz;
      st->rz = tot;
      st->k = st->k + 1;
    }
  }
}

// p = z + beta p  (MODE 0: p = z)
template <int MODE>
__global__ void __launch_bounds__(kBlock)
k_pcg_p(int n, const float* __restrict__ z, float* __restrict__ p, const PcgState* st) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (MODE == 0) p[i] = z[i];
  else p[i] = fmaf((float)st->beta, p[i], z[i]);
}

// ------------------------------------------------------------------------------------------
// Reduced-precision V-cycle (extension). Each level is used in its symmetrically scaled form
// A^ = D^-1/2 A D^-1/2 (unit diagonal, |a^_ij| <= 1 for SPD A, so FP16's range suffices),
// values stored as FP16 or BF16 in the level's SELL positions, products accumulated in FP32.
// Levels whose column offsets j - i fit 16 bits store one packed 32-bit word per entry,
// (j - i) << 16 | half bits, so a lane's four entries arrive in one 16-byte load; the others
// keep the FP32 level's int32 columns beside a 16-bit value array.
// With x = D^-1/2 x^ and b = D^1/2 b^ the weighted-Jacobi V-cycle on A^ is the V-cycle on A:
//   sweep     x^' = x^ + omega (b^ - A^ x^)                 (diag(A^) = 1)
//   restrict  b^_c[I] = isd_c[I] sum_{i in I} sd_i r^_i
//   prolong   x^_i += sd_i isd_c[agg_i] x^_c[agg_i]
// ------------------------------------------------------------------------------------------
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

template <int VT>   // 1: FP16, 2: BF16
__device__ __forceinline__ float h2f(unsigned short u) {
  if (VT == 1) return (float)__builtin_bit_cast(_Float16, u);
  return __uint_as_float((unsigned int)u << 16);
}
template <int VT>
__device__ __forceinline__ unsigned short f2h(float f) {
  if (VT == 1) return __builtin_bit_cast(unsigned short, (_Float16)f);
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}

template <int VT, bool PACKED>
__device__ __forceinline__ float sellh_row(int i, const int* __restrict__ slice_ptr,
                                           const int* __restrict__ idx,
                                           const uint32_t* __restrict__ hpk,
                                           const uint16_t* __restrict__ hval,
                                           const float* __restrict__ x) {
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  float acc = 0.0f;
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    int c0, c1, c2, c3;
    unsigned short h0, h1, h2, h3;
    if (PACKED) {
      const v4u32 e = __builtin_nontemporal_load(reinterpret_cast<const v4u32*>(hpk + p));
      c0 = i + ((int)e.x >> 16); c1 = i + ((int)e.y >> 16);
      c2 = i + ((int)e.z >> 16); c3 = i + ((int)e.w >> 16);
      h0 = (unsigned short)e.x; h1 = (unsigned short)e.y;
      h2 = (unsigned short)e.z; h3 = (unsigned short)e.w;
    } else {
      const v4u16 hv = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(hval + p));
      const v4i c = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(idx + p));
      c0 = c.x; c1 = c.y; c2 = c.z; c3 = c.w;
      h0 = hv.x; h1 = hv.y; h2 = hv.z; h3 = hv.w;
    }
    acc = fmaf(h2f<VT>(h0), x[c0], acc);
    acc = fmaf(h2f<VT>(h1), x[c1], acc);
    acc = fmaf(h2f<VT>(h2), x[c2], acc);
    acc = fmaf(h2f<VT>(h3), x[c3], acc);
  }
  return acc;
}

// MODE 1: y = x + omega (b - A^x);  MODE 2: y = b - A^x;  MODE 3: y = isd (x + omega (b - A^x));
// MODE 4 (Chebyshev step, unit diagonal): y = x + c1 (x - xp) + c2 (b - A^x);  MODE 5: isd times
// the MODE 4 value
template <int VT, bool PACKED, int MODE>
__global__ void __launch_bounds__(kBlock)
k_sellh(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
        const uint32_t* __restrict__ hpk, const uint16_t* __restrict__ hval,
        const float* __restrict__ x, float* y, const float* __restrict__ b,
        float omega, const float* __restrict__ wp, const float* __restrict__ isd, const int* __restrict__ done,
        const float* xp, const float* __restrict__ cc) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const float ax = sellh_row<VT, PACKED>(i, slice_ptr, idx, hpk, hval, x);
  if (i >= n) return;
  float out;
  if (MODE == 2) {
    out = b[i] - ax;
  } else if (MODE == 1 || MODE == 3) {
    const float xn = fmaf(omega, b[i] - ax, x[i]);
    out = MODE == 3 ? isd[i] * xn : xn;
  } else {
    const float xi = x[i];
    const float u = fmaf(cc[0], xi - (xp ? xp[i] : 0.0f), xi);
    const float xn = fmaf(cc[1], b[i] - ax, u);
    out = MODE == 5 ? isd[i] * xn : xn;
  }
  y[i] = out;
}

// Power iteration on the scaled reduced-precision level (unit diagonal, so D = I in these
// variables and the spectrum is that of D^-1 A as the V-cycle applies it): w = A^ v and the
// partials v.A^v (slot 0), v.v (slot 1), w.w (slot 2) for k_power_fin.
template <int VT, bool PACKED>
__global__ void __launch_bounds__(kBlock)
k_power_step_h(int n, int n_slices, const int* __restrict__ slice_ptr, const int* __restrict__ idx,
               const uint32_t* __restrict__ hpk, const uint16_t* __restrict__ hval,
               const float* __restrict__ v, float* __restrict__ w, RedWS ws) {
  double vav = 0.0, vv = 0.0, ww = 0.0;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = sellh_row<VT, PACKED>(i, slice_ptr, idx, hpk, hval, v);
    if (i < n) {
      w[i] = ax;
      vav = (double)v[i] * ax;
      vv = (double)v[i] * v[i];
      ww = (double)ax * ax;
    }
  }
  const double s0 = block_sum(vav);
  const double s1 = block_sum(vv);
  const double s2 = block_sum(ww);
  store_partial(ws, 0, s0);
  store_partial(ws, 1, s1);
  store_partial(ws, 2, s2);
}

// ------------------------------------------------------------------------------------------
// Windowed packed levels: the x values a block of kBlock rows reads are staged in LDS first.
// The block's window is the sorted list of the 32-column segments its rows' columns fall in
// (segs[seg_ptr[blk] ..]), segment k at LDS floats [32 k, 32 k + 32); the high half of each
// packed word holds the LDS byte offset of its column, (32 k + col % 32) * 4, instead of the
// column offset j - i. A row then streams 4 bytes per entry from HBM and reads x from LDS, and
// x is fetched once per block, 128 coalesced bytes per segment, instead of once per entry.
// Sums run over the same words in the same order as the global-gather rows above, so results
// are bit-identical to them. Blocks go to XCDs round-robin (block b on XCD b % 8): row block
// for launch block b keeps each XCD on a contiguous range of rows, so its L2 holds the part
// of x that its windows cover.
// ------------------------------------------------------------------------------------------
constexpr int kWinSeg = MGPBD_WINDOW_SEG;          // columns per segment
constexpr int kWinMaxSegs = MGPBD_WINDOW_MAX_SEGS; // 48 KiB of LDS per block; offsets fit 16 bits
static_assert(MGPBD_WINDOW_ROWS == kBlock, "window blocks are the sweep workgroups");
static_assert(kWinSeg * kWinMaxSegs * 4 <= 65536, "LDS byte offsets fit 16 bits");

__device__ __forceinline__ int xcd_row_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, k = b & 7;
  return k * q + min(k, r) + (b >> 3);
}

// Stages block blk's window of x (n_x entries; the last segment's tail past n_x reads entry
// n_x - 1, which no column refers to) into xs (all threads of the block take part).
__device__ __forceinline__ void stage_window(float* xs, const int* __restrict__ seg_ptr,
                                             const int* __restrict__ segs, int blk,
                                             const float* __restrict__ x, int n_x) {
  const int s0 = seg_ptr[blk];
  const int ns = seg_ptr[blk + 1] - s0;   // >= 1: every block holds a row with a diagonal
  const int sub = threadIdx.x >> 5, ln = threadIdx.x & 31;
  constexpr int kPer = kBlock / kWinSeg;  // segments per pass
  // passes past the window repeat its last segment: same slot, same value
  for (int k0 = 0; k0 < ns; k0 += 8 * kPer) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = min(k0 + kPer * u + sub, ns - 1);
      v[u] = x[min(segs[s0 + k] * kWinSeg + ln, n_x - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = min(k0 + kPer * u + sub, ns - 1);
      xs[k * kWinSeg + ln] = v[u];
    }
  }
  __syncthreads();
}

template <int VT>
__device__ __forceinline__ float sellw_row(int i, const int* __restrict__ slice_ptr,
                                           const uint32_t* __restrict__ hpk, const float* xs) {
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const char* xl = reinterpret_cast<const char*>(xs);
  float acc = 0.0f;
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const v4u32 e = __builtin_nontemporal_load(reinterpret_cast<const v4u32*>(hpk + p0 + 64 * k0));
    acc = fmaf(h2f<VT>((unsigned short)e.x), *reinterpret_cast<const float*>(xl + (e.x >> 16)), acc);
    acc = fmaf(h2f<VT>((unsigned short)e.y), *reinterpret_cast<const float*>(xl + (e.y >> 16)), acc);
    acc = fmaf(h2f<VT>((unsigned short)e.z), *reinterpret_cast<const float*>(xl + (e.z >> 16)), acc);
    acc = fmaf(h2f<VT>((unsigned short)e.w), *reinterpret_cast<const float*>(xl + (e.w >> 16)), acc);
  }
  return acc;
}

// Windowed counterpart of k_sellh (same MODE semantics), dynamic LDS of max_segs * 128 bytes.
template <int VT, int MODE>
__global__ void __launch_bounds__(kBlock)
k_sellw(int n, int n_slices, const int* __restrict__ slice_ptr, const uint32_t* __restrict__ hpk,
        const int* __restrict__ seg_ptr, const int* __restrict__ segs,
        const float* __restrict__ x, float* y, const float* __restrict__ b,
        float omega, const float* __restrict__ wp, const float* __restrict__ isd,
        const int* __restrict__ done, const float* xp, const float* __restrict__ cc) {
  extern __shared__ float xs[];
  if (wp) omega = *wp;
  if (done && *done) return;
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, x, n);
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const float ax = sellw_row<VT>(i, slice_ptr, hpk, xs);
  if (i >= n) return;
  float out;
  if (MODE == 2) {
    out = b[i] - ax;
  } else if (MODE == 1 || MODE == 3) {
    const float xn = fmaf(omega, b[i] - ax, x[i]);
    out = MODE == 3 ? isd[i] * xn : xn;
  } else {
    const float xi = x[i];
    const float u = fmaf(cc[0], xi - (xp ? xp[i] : 0.0f), xi);
    const float xn = fmaf(cc[1], b[i] - ax, u);
    out = MODE == 5 ? isd[i] * xn : xn;
  }
  y[i] = out;
}

// Windowed counterpart of k_power_step_h; partials stored at the row block's slot, so the
// totals are those of k_power_step_h.
template <int VT>
__global__ void __launch_bounds__(kBlock)
k_power_step_w(int n, int n_slices, const int* __restrict__ slice_ptr,
               const uint32_t* __restrict__ hpk, const int* __restrict__ seg_ptr,
               const int* __restrict__ segs, const float* __restrict__ v, float* __restrict__ w,
               RedWS ws) {
  extern __shared__ float xs[];
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, v, n);
  double vav = 0.0, vv = 0.0, ww = 0.0;
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = sellw_row<VT>(i, slice_ptr, hpk, xs);
    if (i < n) {
      w[i] = ax;
      vav = (double)v[i] * ax;
      vv = (double)v[i] * v[i];
      ww = (double)ax * ax;
    }
  }
  const double s0 = block_sum(vav);
  const double s1 = block_sum(vv);
  const double s2 = block_sum(ww);
  if (threadIdx.x == 0) {
    ws.partials[blk] = s0;
    ws.partials[(size_t)kPartialStride + blk] = s1;
    ws.partials[(size_t)2 * kPartialStride + blk] = s2;
  }
}

// LDS byte offset of each slot's column in its block's window (binary search of the sorted
// segment list; 0 for padding rows, whose values are 0): into the high halves of a windowed
// level's packed words (hpk, low halves -- the value bits -- kept) and into the 16-bit offset
// array of the windowed FP32 kernels (woff); either may be NULL.
__global__ void __launch_bounds__(kBlock)
k_window_offsets(int n, int n_slices, const int* __restrict__ slice_ptr,
                 const int* __restrict__ idx, const int* __restrict__ seg_ptr,
                 const int* __restrict__ segs, uint32_t* __restrict__ hpk,
                 uint16_t* __restrict__ woff, int* __restrict__ bad) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const int blk = i / kBlock;
  const int lo0 = seg_ptr[blk], hi0 = seg_ptr[blk + 1];
  for (int k = 0; k < w; ++k) {
    const int p = p0 + 64 * (k & ~3) + (k & 3);
    uint32_t off = 0;
    if (i < n) {
      const int c = idx[p];
      const int sg = c / kWinSeg;
      int lo = lo0, hi = hi0;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid] < sg) lo = mid + 1; else hi = mid;
      }
      if (lo < hi0 && segs[lo] == sg) off = (uint32_t)((lo - lo0) * kWinSeg + c % kWinSeg) * 4u;
      else atomicOr(bad, 1);
    }
    if (hpk) hpk[p] = (off << 16) | (hpk[p] & 0xFFFFu);
    if (woff) woff[p] = (uint16_t)off;
  }
}

// Windowed FP32 rows: FP32 values beside 16-bit LDS offsets (6 bytes per entry), x from the
// block's window in LDS; products summed in the entry order of sell_row / sell_row_d, so the
// results equal theirs bit for bit.
template <bool ACC64>
__device__ __forceinline__ double sellw_row_f32(int i, const int* __restrict__ slice_ptr,
                                                const uint16_t* __restrict__ woff,
                                                const float* __restrict__ val, const float* xs) {
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const char* xl = reinterpret_cast<const char*>(xs);
  double accd = 0.0;
  float accf = 0.0f;
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(val + p));
    const v4u16 o = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(woff + p));
    const float x0 = *reinterpret_cast<const float*>(xl + o.x);
    const float x1 = *reinterpret_cast<const float*>(xl + o.y);
    const float x2 = *reinterpret_cast<const float*>(xl + o.z);
    const float x3 = *reinterpret_cast<const float*>(xl + o.w);
    if (ACC64) {
      accd = fma((double)v.x, (double)x0, accd);
      accd = fma((double)v.y, (double)x1, accd);
      accd = fma((double)v.z, (double)x2, accd);
      accd = fma((double)v.w, (double)x3, accd);
    } else {
      accf = fmaf(v.x, x0, accf);
      accf = fmaf(v.y, x1, accf);
      accf = fmaf(v.z, x2, accf);
      accf = fmaf(v.w, x3, accf);
    }
  }
  return ACC64 ? accd : (double)accf;
}

// Windowed k_pcg_spmv_dot (partials at the row block's slot: the same totals)
template <bool ACC64>
__global__ void __launch_bounds__(kBlock)
k_pcg_spmv_dot_w(int n, int n_slices, const int* __restrict__ slice_ptr,
                 const uint16_t* __restrict__ woff, const float* __restrict__ val,
                 const int* __restrict__ seg_ptr, const int* __restrict__ segs,
                 const float* __restrict__ p, float* __restrict__ Ap, RedWS ws,
                 const PcgState* st) {
  extern __shared__ float xs[];
  if (st->done) return;
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, p, n);
  double acc = 0.0;
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = (float)sellw_row_f32<ACC64>(i, slice_ptr, woff, val, xs);
    if (i < n) {
      Ap[i] = ax;
      acc = (double)p[i] * (double)ax;
    }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) ws.partials[blk] = t;
}

// Windowed k_sell (same MODE semantics and arithmetic, FP32 values beside LDS offsets)
template <int MODE>
__global__ void __launch_bounds__(kBlock)
k_sellw_f32(int n, int n_slices, const int* __restrict__ slice_ptr,
            const uint16_t* __restrict__ woff, const float* __restrict__ val,
            const int* __restrict__ seg_ptr, const int* __restrict__ segs,
            const float* __restrict__ x, float* y, const float* __restrict__ dinv,
            const float* __restrict__ b, float omega, const float* __restrict__ wp,
            const int* __restrict__ done, const float* xp, const float* __restrict__ cc) {
  extern __shared__ float xs[];
  if (wp) omega = *wp;
  if (done && *done) return;
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, x, n);
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const float ax = (float)sellw_row_f32<false>(i, slice_ptr, woff, val, xs);
  if (i >= n) return;
  if (MODE == 0) y[i] = ax;
  else if (MODE == 1) y[i] = fmaf(omega * dinv[i], b[i] - ax, x[i]);
  else if (MODE == 2) y[i] = b[i] - ax;
  else {
    const float xi = x[i];
    const float u = fmaf(cc[0], xi - (xp ? xp[i] : 0.0f), xi);
    y[i] = fmaf(cc[1] * dinv[i], b[i] - ax, u);
  }
}

// Windowed k_power_step (partials at the row block's slot)
__global__ void __launch_bounds__(kBlock)
k_power_step_wf(int n, int n_slices, const int* __restrict__ slice_ptr,
                const uint16_t* __restrict__ woff, const float* __restrict__ val,
                const int* __restrict__ seg_ptr, const int* __restrict__ segs,
                const float* __restrict__ dinv, const float* __restrict__ v,
                float* __restrict__ w, RedWS ws) {
  extern __shared__ float xs[];
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, v, n);
  double vav = 0.0, vdv = 0.0, wdw = 0.0;
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const float ax = (float)sellw_row_f32<false>(i, slice_ptr, woff, val, xs);
    if (i < n) {
      const float di = dinv[i];
      w[i] = di * ax;
      vav = (double)v[i] * ax;
      if (di > 0.0f) {
        vdv = (double)v[i] * v[i] / di;
        wdw = (double)ax * ax * di;
      }
    }
  }
  const double s0 = block_sum(vav);
  const double s1 = block_sum(vdv);
  const double s2 = block_sum(wdw);
  if (threadIdx.x == 0) {
    ws.partials[blk] = s0;
    ws.partials[(size_t)kPartialStride + blk] = s1;
    ws.partials[(size_t)2 * kPartialStride + blk] = s2;
  }
}

// Windowed k_true_resid (partials at the row block's slot)
__global__ void __launch_bounds__(kBlock)
k_true_resid_w(int n, int n_slices, const int* __restrict__ slice_ptr,
               const uint16_t* __restrict__ woff, const float* __restrict__ val,
               const int* __restrict__ seg_ptr, const int* __restrict__ segs,
               const float* __restrict__ x, const float* __restrict__ b, RedWS ws) {
  extern __shared__ float xs[];
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, x, n);
  double acc = 0.0;
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) < n_slices) {
    const double ax = sellw_row_f32<true>(i, slice_ptr, woff, val, xs);
    if (i < n) {
      const double ri = (double)b[i] - ax;
      acc = ri * ri;
    }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) ws.partials[blk] = t;
}

// sd = sqrt(A_ii), isd = 1 / sd (1 where A_ii <= 0)
__global__ void __launch_bounds__(kBlock)
k_half_scales(int n, const int* __restrict__ diag_pos, const float* __restrict__ val,
              float* __restrict__ sd, float* __restrict__ isd) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int p = diag_pos[i];
  const float d = p >= 0 ? val[p] : 0.0f;
  const float r = d > 0.0f ? sqrtf(d) : 1.0f;
  sd[i] = r;
  isd[i] = 1.0f / r;
}

// half(val_ij isd_i isd_j) over the whole padded row, four entries (one 16-byte index and
// value load) at a time, stored as packed words (j - i) << 16 | half (one 16-byte store;
// padding: 0) when hpk != NULL, else into hval (one 8-byte store); with keep_hi (windowed
// levels) the words' high halves, the LDS offsets, are kept and only the values replaced
template <int VT>
__global__ void __launch_bounds__(kBlock)
k_half_values(int n, int n_slices, const int* __restrict__ slice_ptr,
              const int* __restrict__ idx, const float* __restrict__ val,
              const float* __restrict__ isd, uint16_t* __restrict__ hval,
              uint32_t* __restrict__ hpk, int keep_hi) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const bool row = i < n;
  const float si = row ? isd[i] : 0.0f;
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    const v4i c = *reinterpret_cast<const v4i*>(idx + p);
    const v4f v = *reinterpret_cast<const v4f*>(val + p);
    v4u16 h;
    h.x = f2h<VT>(row ? v.x * si * isd[c.x] : 0.0f);
    h.y = f2h<VT>(row ? v.y * si * isd[c.y] : 0.0f);
    h.z = f2h<VT>(row ? v.z * si * isd[c.z] : 0.0f);
    h.w = f2h<VT>(row ? v.w * si * isd[c.w] : 0.0f);
    if (hpk && keep_hi) {   // windowed level: LDS offsets stay
      v4u32 e = *reinterpret_cast<const v4u32*>(hpk + p);
      e.x = (e.x & 0xFFFF0000u) | h.x;
      e.y = (e.y & 0xFFFF0000u) | h.y;
      e.z = (e.z & 0xFFFF0000u) | h.z;
      e.w = (e.w & 0xFFFF0000u) | h.w;
      *reinterpret_cast<v4u32*>(hpk + p) = e;
    } else if (hpk) {
      v4u32 e;
      e.x = (row ? (uint32_t)(c.x - i) << 16 : 0u) | h.x;
      e.y = (row ? (uint32_t)(c.y - i) << 16 : 0u) | h.y;
      e.z = (row ? (uint32_t)(c.z - i) << 16 : 0u) | h.z;
      e.w = (row ? (uint32_t)(c.w - i) << 16 : 0u) | h.w;
      *reinterpret_cast<v4u32*>(hpk + p) = e;
    } else {
      *reinterpret_cast<v4u16*>(hval + p) = h;
    }
  }
}

// k_half_values on a windowed level with FP32 offsets (level 0): isd staged through the
// block's window, each word built from the FP32 value and the slot's LDS offset, so per slot
// 4 + 2 bytes are read and 4 written (no column indices, no old words, no isd gathers); the
// words equal k_half_values' bit for bit (same product order; padding rows: offset 0, value 0).
template <int VT>
__global__ void __launch_bounds__(kBlock)
k_half_values_w(int n, int n_slices, const int* __restrict__ slice_ptr,
                const uint16_t* __restrict__ woff, const float* __restrict__ val,
                const int* __restrict__ seg_ptr, const int* __restrict__ segs,
                const float* __restrict__ isd, uint32_t* __restrict__ hpk) {
  extern __shared__ float xs[];
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  stage_window(xs, seg_ptr, segs, blk, isd, n);
  const int i = blk * kBlock + threadIdx.x;
  if ((i >> 6) >= n_slices) return;
  const int p0 = sell_base(slice_ptr, i);
  const int w = sell_width(slice_ptr, i);
  const bool row = i < n;
  const float si = row ? isd[i] : 0.0f;
  const char* xl = reinterpret_cast<const char*>(xs);
#pragma unroll 4
  for (int k0 = 0; k0 < w; k0 += 4) {
    const int p = p0 + 64 * k0;
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(val + p));
    const v4u16 o = __builtin_nontemporal_load(reinterpret_cast<const v4u16*>(woff + p));
    v4u32 e;
    e.x = ((uint32_t)o.x << 16) | f2h<VT>(row ? v.x * si * *reinterpret_cast<const float*>(xl + o.x) : 0.0f);
    e.y = ((uint32_t)o.y << 16) | f2h<VT>(row ? v.y * si * *reinterpret_cast<const float*>(xl + o.y) : 0.0f);
    e.z = ((uint32_t)o.z << 16) | f2h<VT>(row ? v.z * si * *reinterpret_cast<const float*>(xl + o.z) : 0.0f);
    e.w = ((uint32_t)o.w << 16) | f2h<VT>(row ? v.w * si * *reinterpret_cast<const float*>(xl + o.w) : 0.0f);
    __builtin_nontemporal_store(e, reinterpret_cast<v4u32*>(hpk + p));
  }
}

// x^0 = omega isd r (first sweep on the scaled level 0) and b^0 = isd r
__global__ void __launch_bounds__(kBlock)
k_half_first(int n, const float* __restrict__ isd, const float* __restrict__ r,
             float* __restrict__ bh, float* __restrict__ xh, float omega, const float* __restrict__ wp,
             const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float v = isd[i] * r[i];
  bh[i] = v;
  xh[i] = omega * v;
}

__global__ void __launch_bounds__(kBlock)
k_restrict_h(int nc, const int* __restrict__ mptr, const int* __restrict__ mem,
             const float* __restrict__ sd_f, const float* __restrict__ r,
             const float* __restrict__ isd_c, float* __restrict__ bc, float* __restrict__ xc,
             float omega, const float* __restrict__ wp, const int* __restrict__ done) {
  if (wp) omega = *wp;   // per-level weight held in device memory
  if (done && *done) return;
  const int I = blockIdx.x * kBlock + threadIdx.x;
  if (I >= nc) return;
  const float s = agg_sum<true>(I, mptr, mem, sd_f, r);
  const float v = isd_c[I] * s;
  bc[I] = v;
  if (xc) xc[I] = omega * v;
}

__global__ void __launch_bounds__(kBlock)
k_prolong_h(int n, const int* __restrict__ agg, const float* __restrict__ sd_f,
            const float* __restrict__ isd_c, const float* __restrict__ xc,
            float* __restrict__ x, float g, const int* __restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int a = agg[i];
  x[i] = fmaf((sd_f[i] * isd_c[a]) * g, xc[a], x[i]);
}

// x^ = sd (D (sd b^)) for the dense coarsest inverse D of the unscaled level
__global__ void __launch_bounds__(kBlock)
k_dense_matvec_scaled(int n, const double* __restrict__ D, const float* __restrict__ sd,
                      const float* __restrict__ b, float* __restrict__ x,
                      const int* __restrict__ done, Push push) {
  if (done && *done) return;
  const int i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* col = D + (size_t)i * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += col[j] * ((double)sd[j] * (double)b[j]);
  acc = wave_sum(acc);   // the xor butterfly leaves the same total in every lane
  const float xi = (float)(acc * (double)sd[i]);
  if (lane == 0) x[i] = xi;
  if (push.x) {   // the row's members, one lane each (k_prolong_h's arithmetic)
    const int m1 = push.mptr[i + 1];
    for (int m = push.mptr[i] + lane; m < m1; m += 64) {
      const int j = push.mem[m];
      push.x[j] = fmaf((push.sd_f[j] * push.isd_c[i]) * push.g, xi, push.x[j]);
    }
  }
}

// Flexible PCG (Polak-Ribiere): partial r.z and r.z_old
__global__ void __launch_bounds__(kBlock)
k_pcg_rz2(int n, const float* __restrict__ r, const float* __restrict__ z,
          const float* __restrict__ zo, RedWS ws, const PcgState* st) {
  if (st->done) return;
  double a = 0.0, c = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double ri = r[i];
    a += ri * (double)z[i];
    c += ri * (double)zo[i];
  }
  const double ba = block_sum(a);
  const double bc = block_sum(c);
  store_partial(ws, 0, ba);
  store_partial(ws, 1, bc);
}

// beta = (r.z - r.z_old) / rz_prev; breakdown test on rz_prev as in the reference
__global__ void __launch_bounds__(kBlock)
k_pcg_fin_flex(RedWS ws, int nparts, PcgState* st) {
  if (st->done) return;
  double t[2];
  sum_partials_k<2>(ws, 0, nparts, t);
  const double rz = t[0], rzo = t[1];
  if (threadIdx.x != 0) return;
  st->iters = st->k + 1;
  if (fabs(st->rz) < 1e-30) {
    st->done = 3;
  } else {
    st->beta = (rz - rzo) / st->rz;
    st->rz = rz;
    st->k = st->k + 1;
  }
}

// p = z + beta p and z_old = z
__global__ void __launch_bounds__(kBlock)
k_pcg_p_flex(int n, const float* __restrict__ z, float* __restrict__ p, float* __restrict__ zo,
             const PcgState* st, int first) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float zi = z[i];
  p[i] = first ? zi : fmaf((float)st->beta, p[i], zi);
  zo[i] = zi;
}

// ------------------------------------------------------------------------------------------
// Factored level 0 of the reduced-precision V-cycle. The level-0 matrix is
// A = J W J^T + diag(alpha~): row t of J holds tet t's vertex gradients g_tk at its vertices'
// columns, W = diag(inverse masses). Its scaled form D^-1/2 A D^-1/2 is therefore
// K W' K^T + diag(a^) with K_tk = c g_tk / sqrt(A_tt), W' = W / c^2 and a^_t = alpha~_t / A_tt,
// c^2 = the largest inverse mass (so |K_tk| = sqrt(w_max / w_v) times what the unit-weight
// form sqrt(w_v) g_tk / sqrt(A_tt) <= 1 holds: ~11 bits of every component in FP16).
// A tet's constraint is a function of its deformation gradient, so its four vertex gradients
// sum to zero: the tet order stores the first three rounded, and the fourth is defined as
// the rounded negative sum of those three (k_fac_k3) in both copies. y = A^ x takes two
// passes:
//   per vertex  u_v = w'_v sum over the tets t holding v of K_tk x_t   (k_fac_gather)
//   per tet     y_t = a^_t x_t + sum_k K_tk . u_v(t,k)                  (k_fac_apply)
// The vertex pass reads a vertex-sliced incidence list (vs_inc = 4 tet + local vertex) with K
// beside it (four halves per slot, the fourth zero): four lanes per vertex, 16 vertices per
// 64-lane slice, slot k of vertex 16s+l at vs_ptr[s] + 64 (k/4) + 4l + k%4, so lane q of a
// vertex takes its slots q, q+4, ... with coalesced loads across the wavefront and the four
// partial sums meet in two shuffles (one lane per vertex left too few wavefronts to cover the
// gathers' latency). The tet pass reads K in tet order: nine halves per tet, eight in a
// 16-byte plane and the ninth in a 2-byte plane (FacK). Both copies hold the same bits, so
// the operator is exactly symmetric. Per application this streams ~12 bytes per incidence
// slot and ~42 per tet (~115 MB at 1M tets) instead of ~4 bytes per stored entry of the
// assembled half level (~307 MB), and the per-Newton-iteration conversion reads the
// gradients instead of the 72.5M FP32 entries.
// ------------------------------------------------------------------------------------------
typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));
// Per-vertex products u of the vertex pass: three floats, 12 bytes (one dwordx3 access)
struct fvec3 {
  float x, y, z;
};

template <int VT>
__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
  return (uint32_t)f2h<VT>(a) | ((uint32_t)f2h<VT>(b) << 16);
}
template <int VT>
__device__ __forceinline__ float lo_h(uint32_t w) { return h2f<VT>((unsigned short)w); }
template <int VT>
__device__ __forceinline__ float hi_h(uint32_t w) { return h2f<VT>((unsigned short)(w >> 16)); }

// A K component in the storage format; FP16 saturates at +-65504 (a tet whose vertices are
// almost all pinned can scale a pinned vertex's gradient far up, and an infinity there would
// turn that vertex's zero weight into NaN).
template <int VT>
__device__ __forceinline__ unsigned short f2h_sat(float f) {
  if (VT == 1) f = fminf(fmaxf(f, -65504.0f), 65504.0f);
  return f2h<VT>(f);
}

// K of a level-0 tet in tet order: halves 0..7 (g_0 xyz, g_1 xyz, g_2 xy) in one 16-byte
// plane, half 8 (g_2 z) in a 2-byte plane.
struct FacK {
  const v4u32* __restrict__ a;
  const unsigned short* __restrict__ z;
};

// One component of the fourth vertex's K from the other three (stored) ones: the rounded
// negative sum, in a fixed order, so every reader gets the same bits.
template <int VT>
__device__ __forceinline__ unsigned short k_fac_k3(float k0, float k1, float k2) {
  return f2h_sat<VT>(-((k0 + k1) + k2));
}

// The level-0 diagonal A_tt = sum_k w_k |g_tk|^2 + alpha~_t from the gradients (sd, isd, 1/A_tt),
// then K = cs isd_t g_tk of the first three vertices in tet order (FacK) and a^; k_fac_vert
// copies the same K bits (and the fourth vertex's, k_fac_k3) into vertex-slot order.
template <int VT>
__global__ void __launch_bounds__(kBlock)
k_fac_tet(int n, const int4* __restrict__ t_i, const float* __restrict__ invm,
          const float* __restrict__ gradC, const float* __restrict__ alpha, float cs,
          float* __restrict__ sd, float* __restrict__ isd, float* __restrict__ dinv,
          v4u32* __restrict__ tka, unsigned short* __restrict__ tkz, float* __restrict__ ahat) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int4 vi = t_i[t];
  const float4* g4 = reinterpret_cast<const float4*>(gradC + 12 * (size_t)t);
  const float4 a = g4[0], b = g4[1], c = g4[2];
  // g_0 = (a.x a.y a.z), g_1 = (a.w b.x b.y), g_2 = (b.z b.w c.x), g_3 = (c.y c.z c.w)
  const float w0 = invm[vi.x], w1 = invm[vi.y], w2 = invm[vi.z], w3 = invm[vi.w];
  float d = alpha[t];
  d = fmaf(w0, fmaf(a.x, a.x, fmaf(a.y, a.y, a.z * a.z)), d);
  d = fmaf(w1, fmaf(a.w, a.w, fmaf(b.x, b.x, b.y * b.y)), d);
  d = fmaf(w2, fmaf(b.z, b.z, fmaf(b.w, b.w, c.x * c.x)), d);
  d = fmaf(w3, fmaf(c.y, c.y, fmaf(c.z, c.z, c.w * c.w)), d);
  const float r = d > 0.0f ? sqrtf(d) : 1.0f;
  const float is = 1.0f / r;
  sd[t] = r;
  isd[t] = is;
  dinv[t] = d > 0.0f ? 1.0f / d : 0.0f;
  const float s = cs * is;
  const auto h2 = [&](float p, float q) {
    return (uint32_t)f2h_sat<VT>(s * p) | ((uint32_t)f2h_sat<VT>(s * q) << 16);
  };
  tka[t] = v4u32{h2(a.x, a.y), h2(a.z, a.w), h2(b.x, b.y), h2(b.z, b.w)};
  tkz[t] = f2h_sat<VT>(s * c.x);
  ahat[t] = (alpha[t] * is) * is;
}

// The four vertices' K of tet t from the tet-order planes (the fourth by k_fac_k3), FP32.
template <int VT>
__device__ __forceinline__ void fac_tet_k(FacK tk, int t, float (&k)[12]) {
  const v4u32 q = __builtin_nontemporal_load(tk.a + t);
  const unsigned short z = __builtin_nontemporal_load(tk.z + t);
  k[0] = lo_h<VT>(q.x); k[1] = hi_h<VT>(q.x); k[2] = lo_h<VT>(q.y);
  k[3] = hi_h<VT>(q.y); k[4] = lo_h<VT>(q.z); k[5] = hi_h<VT>(q.z);
  k[6] = lo_h<VT>(q.w); k[7] = hi_h<VT>(q.w); k[8] = h2f<VT>(z);
#pragma unroll
  for (int c = 0; c < 3; ++c) k[9 + c] = h2f<VT>(k_fac_k3<VT>(k[c], k[3 + c], k[6 + c]));
}

constexpr int kFacLanes = 4;   // lanes per vertex in the vertex pass

// First own slot of lane q of vertex v and the number of its own slots (k = q, q + 4, ...).
__device__ __forceinline__ int fac_slot0(const int* __restrict__ vs_ptr, int v, int q) {
  return vs_ptr[v >> 4] + ((v & 15) << 2) + q;
}

// With tbase (mgpbd_mg_set_level0_offsets) the fourth 16-bit word of every K slot holds the
// slot's tet minus the vertex's base tet (raw bits, not a half value) and the fourth float
// of its FP32 gradient the same offset as integer bits: the vertex passes then take the tet
// from the value they load anyway and never read the incidence list.
template <int VT>
__global__ void __launch_bounds__(kBlock)
k_fac_vert(int nv, const int* __restrict__ vs_ptr, const int* __restrict__ vs_cnt,
           const int* __restrict__ vs_inc, const int* __restrict__ tbase,
           const float* __restrict__ gradC, FacK tk, v2u32* __restrict__ vs_k,
           float4* __restrict__ vs_g, const float* __restrict__ invm, float ic2,
           int2* __restrict__ tbw) {
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;
  const int base = fac_slot0(vs_ptr, v, q);
  const int own = (vs_cnt[v] - q + kFacLanes - 1) / kFacLanes;
  const int tb = tbase ? tbase[v] : 0;
  if (tbw && q == 0) tbw[v] = make_int2(tb, __float_as_int(invm[v] * ic2));
  for (int k = 0; k < own; ++k) {
    const int p = base + 64 * k;
    const int inc = vs_inc[p];
    const int t = inc >> 2, j = inc & 3;
    const uint32_t off = tbase ? (uint32_t)(t - tb) : 0u;
    float kt[12];
    fac_tet_k<VT>(tk, t, kt);
    const float* g = gradC + 12 * (size_t)t + 3 * j;
    // re-rounding a value decoded from the storage format returns its bits
    vs_k[p] = v2u32{pack_h2<VT>(kt[3 * j], kt[3 * j + 1]),
                    (uint32_t)f2h<VT>(kt[3 * j + 2]) | (off << 16)};
    vs_g[p] = make_float4(g[0], g[1], g[2], __uint_as_float(off));
  }
}

// Tet of a vertex slot: from the incidence list, or (OFF) base + the offset word of its K / g.
template <bool OFF>
__device__ __forceinline__ int slot_tet(const int* __restrict__ vs_inc, int p, int tb,
                                        uint32_t word) {
  return OFF ? tb + (int)(word >> 16) : __builtin_nontemporal_load(vs_inc + p) >> 2;
}

// u_v = w'_v sum_k K_{t_k, v} x_{t_k} over the vertex's incidence slots (FP32 accumulation): each
// of the vertex's four lanes sums its own slots (loads of up to six issued before their
// products: six slots per lane is an interior vertex of a Kuhn-cube mesh, and a batch longer
// than the lane's slots costs reloads), then lanes q and q^1, q and q^2 add (fixed order,
// deterministic). With slot offsets (tbase) the tets come from the K words themselves.
constexpr int kFacBatch = 6;
template <int VT, bool OFF>
__device__ __forceinline__ void fac_gather_lane(int v, int q, const int* __restrict__ vs_ptr,
                                                const int* __restrict__ vs_cnt,
                                                const int* __restrict__ vs_inc, int tb,
                                                const v2u32* __restrict__ vs_k,
                                                const float* __restrict__ x, float w,
                                                fvec3* __restrict__ u) {
  const int base = fac_slot0(vs_ptr, v, q);
  const int own = (vs_cnt[v] - q + kFacLanes - 1) / kFacLanes;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (int k0 = 0; k0 < own; k0 += kFacBatch) {
    int t[kFacBatch];
    v2u32 h[kFacBatch];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) {
      const int p = base + 64 * min(k0 + j, own - 1);   // past the end: repeats the last slot
      h[j] = __builtin_nontemporal_load(vs_k + p);
      t[j] = slot_tet<OFF>(vs_inc, p, tb, h[j].y);
    }
    float xv[kFacBatch];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) xv[j] = x[t[j]];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) {
      if (k0 + j < own) {
        ax = fmaf(lo_h<VT>(h[j].x), xv[j], ax);
        ay = fmaf(hi_h<VT>(h[j].x), xv[j], ay);
        az = fmaf(lo_h<VT>(h[j].y), xv[j], az);
      }
    }
  }
  ax += __shfl_xor(ax, 1, 64); ay += __shfl_xor(ay, 1, 64); az += __shfl_xor(az, 1, 64);
  ax += __shfl_xor(ax, 2, 64); ay += __shfl_xor(ay, 2, 64); az += __shfl_xor(az, 2, 64);
  if (q == 0) u[v] = fvec3{w * ax, w * ay, w * az};
}

template <int VT>
__global__ void __launch_bounds__(kBlock)
k_fac_gather(int nv, const int* __restrict__ vs_ptr, const int* __restrict__ vs_cnt,
             const int* __restrict__ vs_inc, const int* __restrict__ tbase,
             const v2u32* __restrict__ vs_k, const float* __restrict__ invm, float ic2,
             const float* __restrict__ x, fvec3* __restrict__ u,
             const int* __restrict__ done) {
  if (done && *done) return;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;   // all four lanes of a vertex leave together
  const float w = invm[v] * ic2;
  if (tbase)
    fac_gather_lane<VT, true>(v, q, vs_ptr, vs_cnt, vs_inc, tbase[v], vs_k, x, w, u);
  else
    fac_gather_lane<VT, false>(v, q, vs_ptr, vs_cnt, vs_inc, 0, vs_k, x, w, u);
}

// The vertex pass on the uniform layout: every slice R rows tall (slot k of vertex v at
// (v / 16) 64 R + 64 (k / 4) + 4 (v % 16) + k % 4) and the tets in the K words' offsets, so a
// lane issues its R K-word loads and the load of the vertex's {base tet, w'} word at once,
// with no slice pointer or count to wait for. Slots past a vertex's count hold K = 0 and offset 0 (never written), so
// they add exact zeros after the lane's own slots: the sums equal k_fac_gather's bit for bit.
template <int VT, int R>
__global__ void __launch_bounds__(kBlock)
k_fac_gather_u(int nv, const int2* __restrict__ tbw, const v2u32* __restrict__ vs_k,
               const float* __restrict__ x, fvec3* __restrict__ u,
               const int* __restrict__ done) {
  if (done && *done) return;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;
  const int base = (v >> 4) * (64 * R) + ((v & 15) << 2) + q;
  const int2 tw = tbw[v];   // base tet, w' = invm_v / c^2
  const int tb = tw.x;
  const float w = __int_as_float(tw.y);
  v2u32 h[R];
#pragma unroll
  for (int j = 0; j < R; ++j) h[j] = __builtin_nontemporal_load(vs_k + base + 64 * j);
  float xv[R];
#pragma unroll
  for (int j = 0; j < R; ++j) xv[j] = x[tb + (int)(h[j].y >> 16)];
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    ax = fmaf(lo_h<VT>(h[j].x), xv[j], ax);
    ay = fmaf(hi_h<VT>(h[j].x), xv[j], ay);
    az = fmaf(lo_h<VT>(h[j].y), xv[j], az);
  }
  ax += __shfl_xor(ax, 1, 64); ay += __shfl_xor(ay, 1, 64); az += __shfl_xor(az, 1, 64);
  ax += __shfl_xor(ax, 2, 64); ay += __shfl_xor(ay, 2, 64); az += __shfl_xor(az, 2, 64);
  if (q == 0) u[v] = fvec3{w * ax, w * ay, w * az};
}

// The four vertices of a level-0 row for the tet passes: from the int4 list, or from one
// packed 64-bit word (mgpbd_mg_set_level0_packed_rows): the first vertex in the low bb bits,
// then the other three as signed ob-bit differences from it (8 bytes per row instead of 16).
struct RowsI4 {
  const int4* __restrict__ p;
  __device__ __forceinline__ int4 operator()(int t) const { return p[t]; }
};
struct RowsPk {
  const uint64_t* __restrict__ p;
  int bb, ob;
  __device__ __forceinline__ int4 operator()(int t) const {
    const uint64_t w = p[t];
    const int v0 = (int)(w & ((1ull << bb) - 1));
    // field k (k = 1..3) at bits [bb + (k-1) ob, bb + k ob): shifted to the top, then
    // arithmetic-shifted down to sign-extend it
    const int d1 = (int)((int64_t)(w << (64 - bb - ob)) >> (64 - ob));
    const int d2 = (int)((int64_t)(w << (64 - bb - 2 * ob)) >> (64 - ob));
    const int d3 = (int)((int64_t)(w << (64 - bb - 3 * ob)) >> (64 - ob));
    return make_int4(v0, v0 + d1, v0 + d2, v0 + d3);
  }
};

// (A^ x)_t from u: a^_t x_t + K_t0 . u_v0 + ... + K_t3 . u_v3 (FP32)
template <int VT, class RS>
__device__ __forceinline__ float fac_row(int t, RS rows, FacK tk, const float* __restrict__ ahat,
                                         const fvec3* __restrict__ u, float xt) {
  const int4 vi = rows(t);
  float k[12];
  fac_tet_k<VT>(tk, t, k);
  const fvec3 u0 = u[vi.x], u1 = u[vi.y], u2 = u[vi.z], u3 = u[vi.w];
  float acc = ahat[t] * xt;
  acc = fmaf(k[0], u0.x, acc);
  acc = fmaf(k[1], u0.y, acc);
  acc = fmaf(k[2], u0.z, acc);
  acc = fmaf(k[3], u1.x, acc);
  acc = fmaf(k[4], u1.y, acc);
  acc = fmaf(k[5], u1.z, acc);
  acc = fmaf(k[6], u2.x, acc);
  acc = fmaf(k[7], u2.y, acc);
  acc = fmaf(k[8], u2.z, acc);
  acc = fmaf(k[9], u3.x, acc);
  acc = fmaf(k[10], u3.y, acc);
  acc = fmaf(k[11], u3.z, acc);
  return acc;
}

// Factored counterpart of k_sellw on level 0 (same MODE semantics, rows = tets).
template <int VT, int MODE, class RS>
__global__ void __launch_bounds__(kBlock)
k_fac_apply(int n, RS rows, FacK tk, const float* __restrict__ ahat,
            const fvec3* __restrict__ u, const float* __restrict__ x, float* y,
            const float* __restrict__ b, float omega, const float* __restrict__ wp, const float* __restrict__ isd,
            const int* __restrict__ done, const float* xp, const float* __restrict__ cc) {
  if (wp) omega = *wp;
  if (done && *done) return;
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float xi = x[i];
  const float ax = fac_row<VT>(i, rows, tk, ahat, u, xi);
  if (MODE == 2) {
    y[i] = b[i] - ax;
  } else if (MODE == 1 || MODE == 3) {
    const float xn = fmaf(omega, b[i] - ax, xi);
    y[i] = MODE == 3 ? isd[i] * xn : xn;
  } else {
    const float uu = fmaf(cc[0], xi - (xp ? xp[i] : 0.0f), xi);
    const float xn = fmaf(cc[1], b[i] - ax, uu);
    y[i] = MODE == 5 ? isd[i] * xn : xn;
  }
}

// Power step on the factored level (w = A^ v after k_fac_gather of v) with the partials
// v.A^v, v.v, w.w of k_power_step_h; the grid covers at least n rows.
template <int VT, class RS>
__global__ void __launch_bounds__(kBlock)
k_fac_power(int n, RS rows, FacK tk, const float* __restrict__ ahat,
            const fvec3* __restrict__ u, const float* __restrict__ v, float* __restrict__ w,
            RedWS ws) {
  double vav = 0.0, vv = 0.0, ww = 0.0;
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < n) {
    const float vi = v[i];
    const float ax = fac_row<VT>(i, rows, tk, ahat, u, vi);
    w[i] = ax;
    vav = (double)vi * ax;
    vv = (double)vi * vi;
    ww = (double)ax * ax;
  }
  const double s0 = block_sum(vav);
  const double s1 = block_sum(vv);
  const double s2 = block_sum(ww);
  store_partial(ws, 0, s0);
  store_partial(ws, 1, s1);
  store_partial(ws, 2, s2);
}

// ------------------------------------------------------------------------------------------
// The PCG's operator in factored form: A x = J W J^T x + alpha~ x with FP64 products and sums
// of the FP32 gradients (the exact operator of the FP32 gradients; the assembled FP32 entries
// of cons/mgpbd_gpu.py:273-309 round each A_ij once). Vertex pass u_v = w_v sum_t g_tv x_t
// (four lanes per vertex over the FP32 vertex-order gradients vs_g), tet pass
// (A x)_t = alpha~_t x_t + sum_k g_tk . u_v(t,k). With it the solve, its guessed-start and
// restart residuals and the FP64 check all use one operator and level 0 is never assembled
// per Newton iteration.
// ------------------------------------------------------------------------------------------
// gate 0: skip when *done is set (PCG launches past convergence); gate 1: run only when the
// solve converged and is not yet verified (st: the verification tail of a replay).
template <typename XT, bool OFF>
__device__ __forceinline__ void fac_ogather_lane(int v, int q, const int* __restrict__ vs_ptr,
                                                 const int* __restrict__ vs_cnt,
                                                 const int* __restrict__ vs_inc, int tb,
                                                 const float4* __restrict__ vs_g,
                                                 const float* __restrict__ invm,
                                                 const XT* __restrict__ x,
                                                 double4* __restrict__ u) {
  const int base = fac_slot0(vs_ptr, v, q);
  const int own = (vs_cnt[v] - q + kFacLanes - 1) / kFacLanes;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int k0 = 0; k0 < own; k0 += kFacBatch) {
    int t[kFacBatch];
    float4 g[kFacBatch];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) {
      const int p = base + 64 * min(k0 + j, own - 1);
      g[j] = vs_g[p];
      t[j] = slot_tet<OFF>(vs_inc, p, tb, __float_as_uint(g[j].w) << 16);
    }
    double xv[kFacBatch];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) xv[j] = (double)x[t[j]];
#pragma unroll
    for (int j = 0; j < kFacBatch; ++j) {
      if (k0 + j < own) {
        ax = fma((double)g[j].x, xv[j], ax);
        ay = fma((double)g[j].y, xv[j], ay);
        az = fma((double)g[j].z, xv[j], az);
      }
    }
  }
  ax += __shfl_xor(ax, 1, 64); ay += __shfl_xor(ay, 1, 64); az += __shfl_xor(az, 1, 64);
  ax += __shfl_xor(ax, 2, 64); ay += __shfl_xor(ay, 2, 64); az += __shfl_xor(az, 2, 64);
  if (q == 0) {
    const double w = (double)invm[v];
    u[v] = make_double4(w * ax, w * ay, w * az, 0.0);
  }
}

template <typename XT>
__global__ void __launch_bounds__(kBlock)
k_fac_ogather(int nv, const int* __restrict__ vs_ptr, const int* __restrict__ vs_cnt,
              const int* __restrict__ vs_inc, const int* __restrict__ tbase,
              const float4* __restrict__ vs_g, const float* __restrict__ invm,
              const XT* __restrict__ x, double4* __restrict__ u, const int* __restrict__ done,
              const PcgState* __restrict__ gate_st) {
  if (done && *done) return;
  if (gate_st && (gate_st->done != 1 || gate_st->verified)) return;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;   // all four lanes of a vertex leave together
  if (tbase)
    fac_ogather_lane<XT, true>(v, q, vs_ptr, vs_cnt, vs_inc, tbase[v], vs_g, invm, x, u);
  else
    fac_ogather_lane<XT, false>(v, q, vs_ptr, vs_cnt, vs_inc, 0, vs_g, invm, x, u);
}

// k_fac_ogather on the uniform layout (see k_fac_gather_u; padded slots hold zero gradients
// and offset 0, so the FP64 sums are the per-slice kernel's bit for bit).
template <typename XT, int R>
__global__ void __launch_bounds__(kBlock)
k_fac_ogather_u(int nv, const int* __restrict__ tbase, const float4* __restrict__ vs_g,
                const float* __restrict__ invm, const XT* __restrict__ x,
                double4* __restrict__ u, const int* __restrict__ done,
                const PcgState* __restrict__ gate_st) {
  if (done && *done) return;
  if (gate_st && (gate_st->done != 1 || gate_st->verified)) return;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;
  const int base = (v >> 4) * (64 * R) + ((v & 15) << 2) + q;
  const int tb = tbase[v];
  float4 g[R];
#pragma unroll
  for (int j = 0; j < R; ++j) g[j] = vs_g[base + 64 * j];
  double xv[R];
#pragma unroll
  for (int j = 0; j < R; ++j) xv[j] = (double)x[tb + (int)__float_as_uint(g[j].w)];
  double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    ax = fma((double)g[j].x, xv[j], ax);
    ay = fma((double)g[j].y, xv[j], ay);
    az = fma((double)g[j].z, xv[j], az);
  }
  ax += __shfl_xor(ax, 1, 64); ay += __shfl_xor(ay, 1, 64); az += __shfl_xor(az, 1, 64);
  ax += __shfl_xor(ax, 2, 64); ay += __shfl_xor(ay, 2, 64); az += __shfl_xor(az, 2, 64);
  if (q == 0) {
    const double w = (double)invm[v];
    u[v] = make_double4(w * ax, w * ay, w * az, 0.0);
  }
}

// dpos_v = w_v sum_t dlambda_t g_tv over the vertex's incidences (FP32, four lanes per vertex
// over the vertex-order gradients, then two shuffles) and lambda += dlambda: the factored
// context's counterpart of k_scatter (same quantity, another summation order).
__global__ void __launch_bounds__(kBlock)
k_fac_scatter(int nv, const int* __restrict__ vs_ptr, const int* __restrict__ vs_cnt,
              const int* __restrict__ vs_inc, const float4* __restrict__ vs_g,
              const float* __restrict__ invm, const float* __restrict__ dlam,
              float* __restrict__ dpos) {
  const int gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int v = gid >> 2, q = gid & 3;
  if (v >= nv) return;
  const int base = fac_slot0(vs_ptr, v, q);
  const int own = (vs_cnt[v] - q + kFacLanes - 1) / kFacLanes;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (int k0 = 0; k0 < own; k0 += 8) {
    int t[8];
    float4 g[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = base + 64 * min(k0 + j, own - 1);
      t[j] = slot_tet<false>(vs_inc, p, 0, 0u);
      g[j] = vs_g[p];
    }
    float dl[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) dl[j] = dlam[t[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (k0 + j < own) {
        ax = fmaf(g[j].x, dl[j], ax);
        ay = fmaf(g[j].y, dl[j], ay);
        az = fmaf(g[j].z, dl[j], az);
      }
    }
  }
  ax += __shfl_xor(ax, 1, 64); ay += __shfl_xor(ay, 1, 64); az += __shfl_xor(az, 1, 64);
  ax += __shfl_xor(ax, 2, 64); ay += __shfl_xor(ay, 2, 64); az += __shfl_xor(az, 2, 64);
  if (q == 0) {
    const float w = invm[v];
    dpos[3 * (size_t)v + 0] = w * ax;
    dpos[3 * (size_t)v + 1] = w * ay;
    dpos[3 * (size_t)v + 2] = w * az;
  }
}

template <class RS>
__device__ __forceinline__ double fac_orow(int t, RS rows, const float* __restrict__ gradC,
                                           const double4* __restrict__ u, double acc) {
  const int4 vi = rows(t);
  const float4* g4 = reinterpret_cast<const float4*>(gradC + 12 * (size_t)t);
  const float4 a = g4[0], b = g4[1], c = g4[2];
  const double4 u0 = u[vi.x], u1 = u[vi.y], u2 = u[vi.z], u3 = u[vi.w];
  acc = fma((double)a.x, u0.x, acc);
  acc = fma((double)a.y, u0.y, acc);
  acc = fma((double)a.z, u0.z, acc);
  acc = fma((double)a.w, u1.x, acc);
  acc = fma((double)b.x, u1.y, acc);
  acc = fma((double)b.y, u1.z, acc);
  acc = fma((double)b.z, u2.x, acc);
  acc = fma((double)b.w, u2.y, acc);
  acc = fma((double)c.x, u2.z, acc);
  acc = fma((double)c.y, u3.x, acc);
  acc = fma((double)c.z, u3.y, acc);
  acc = fma((double)c.w, u3.z, acc);
  return acc;
}

// MODE 0: y = A x rounded to FP32 and the partial x.y (k_pcg_spmv_dot's outputs); MODE 1:
// y = b - A x rounded once; MODE 2: the partial (b - A x)^2 (k_true_resid's). The grid covers
// at least n rows; every block stores its partial (modes 0 and 2).
template <typename XT, int MODE, class RS>
__global__ void __launch_bounds__(kBlock)
k_fac_oapply(int n, RS rows, const float* __restrict__ gradC,
             const float* __restrict__ alpha, const double4* __restrict__ u,
             const XT* __restrict__ x, const float* __restrict__ b, float* __restrict__ y,
             RedWS ws, const int* __restrict__ done, const PcgState* __restrict__ gate_st) {
  if (done && *done) return;
  if (gate_st && (gate_st->done != 1 || gate_st->verified)) return;
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double part = 0.0;
  if (i < n) {
    const double xi = (double)x[i];
    const double ax = fac_orow(i, rows, gradC, u, (double)alpha[i] * xi);
    if (MODE == 0) {
      const float af = (float)ax;
      y[i] = af;
      part = xi * (double)af;
    } else if (MODE == 1) {
      y[i] = (float)((double)b[i] - ax);
    } else {
      const double ri = (double)b[i] - ax;
      part = ri * ri;
    }
  }
  if (MODE != 1) store_partial(ws, 0, block_sum(part));
}

// ------------------------------------------------------------------------------------------
// Level 1 from the gradients. With P[i, agg[i]] = 1, A1 = P^T A P = G W G^T + P^T diag(alpha~) P
// where G_(I,v) = sum of g_tv over the tets t of aggregate I holding vertex v. Per (aggregate,
// vertex) pair: G and w_v (k_fac_gsum); per level-1 entry (I, J) with I <= J: the sum over
// the vertices shared by I and J of w_v G_(I,v) . G_(J,v), plus sum_{i in I} alpha~_i on the
// diagonal (k_fac_galerkin, FP64 sums rounded once, written at the entry's SELL position and
// at its transpose's, so A1 is exactly symmetric). This reads ~2 gathered 16-byte G per shared
// vertex for half the entries instead of the ~13 level-0 entries each level-1 entry sums in
// the refill (cons/mgpcg_gpu.py:183-212 contribution maps).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_fac_gsum(int ng, const int* __restrict__ gptr, const int* __restrict__ gmem,
           const int* __restrict__ gvert, const float* __restrict__ gradC,
           const float* __restrict__ invm, float4* __restrict__ G) {
  const int gp = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gp >= ng) return;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int m = gptr[gp]; m < gptr[gp + 1]; ++m) {
    const int inc = gmem[m];
    const float* g = gradC + 12 * (size_t)(inc >> 2) + 3 * (inc & 3);
    sx += g[0];
    sy += g[1];
    sz += g[2];
  }
  G[gp] = make_float4(sx, sy, sz, invm[gvert[gp]]);
}

__global__ void __launch_bounds__(kBlock)
k_fac_galerkin(int ne, const int* __restrict__ tptr, const int2* __restrict__ terms,
               const float4* __restrict__ G, const int* __restrict__ cpos,
               const int* __restrict__ cpos_t, const int* __restrict__ cagg,
               const int* __restrict__ mptr,
               const int* __restrict__ mem, const float* __restrict__ alpha,
               float* __restrict__ cval) {
  const int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (c >= ne) return;
  double acc = 0.0;
  for (int m = tptr[c]; m < tptr[c + 1]; ++m) {
    const int2 ab = terms[m];
    const float4 ga = G[ab.x], gb = G[ab.y];
    const double d = fma((double)ga.x, (double)gb.x,
                         fma((double)ga.y, (double)gb.y, (double)ga.z * (double)gb.z));
    acc = fma((double)ga.w, d, acc);
  }
  const int I = cagg[c];
  if (I >= 0)
    for (int m = mptr[I]; m < mptr[I + 1]; ++m) acc += (double)alpha[mem[m]];
  const float v = (float)acc;
  cval[cpos[c]] = v;
  cval[cpos_t[c]] = v;   // the transposed entry (the same slot on the diagonal)
}

// ------------------------------------------------------------------------------------------
// Allocation-free constraint pass: the FP64 block partials of ||rhs||^2 live in this
// file-scope array instead of a context's workspace. k_constraints runs at most
// kMaxRedBlocks blocks (grid-stride), so kMaxRedBlocks partials always suffice.
// ------------------------------------------------------------------------------------------
constexpr int kNoAllocPartials = kMaxRedBlocks;               // 2048 doubles = 16 KiB per GPU
constexpr int64_t kNoAllocMaxTets = int64_t(1) << 28;         // 12 * n_tets floats fit size_t math
__device__ __align__(256) double scratch[kNoAllocPartials];

}  // namespace mgpbd

using namespace mgpbd;

namespace {
// Device address of `scratch` on each device, looked up once per device with that device
// current (every GPU holds its own copy of the array at its own address). The host-side
// symbol is only a stand-in: kernels and copies must use these addresses.
constexpr int kMaxScratchDevices = 64;
double* g_scratch_dev[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

int scratch_address(double** out) {
  int dev = -1;
  MGPBD_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints_noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(mgpbd::scratch));
    if (e != hipSuccess || !p)
      return mgpbd_fail(MGPBD_EHIP, std::string("mgpbd_constraints_noalloc: hipGetSymbolAddress: ") +
                                        hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

// ============================================================================================
// Context
// ============================================================================================
struct MgLevel {
  int n = 0, n_slices = 0;
  const int* slice_ptr = nullptr;
  const int* idx = nullptr;
  const float* val = nullptr;
  const float* dinv = nullptr;
  const int* agg = nullptr;
  const int* mptr = nullptr;
  const int* mem = nullptr;
  float* work = nullptr;   // b, x, x', r, two W-cycle and two power-iteration vectors, each
                           // 64 * n_slices long (padded rows addressable)
  int diagonal = 0;
  // reduced-precision copy for the V-cycle (MGPBD_PREC_FP16 / _BF16): values of the scaled
  // level D^-1/2 A D^-1/2 in the SELL positions, packed with 16-bit column offsets (hpk) or
  // beside the int32 columns (hval)
  uint16_t* hval = nullptr;
  uint32_t* hpk = nullptr;
  // LDS window of each block of kBlock rows (windowed packed levels, else null)
  const int* seg_ptr = nullptr;
  const int* segs = nullptr;
  int win_segs = 0;         // largest window (segments) of the level
  uint16_t* woff = nullptr; // LDS offsets beside the FP32 values (windowed FP32 kernels)
  float* sd = nullptr;      // sqrt(A_ii)
  float* isd = nullptr;     // 1 / sqrt(A_ii)
  const int* diag_pos = nullptr;
  float* wb(int k) const { return work + (size_t)k * 64 * n_slices; }
};


struct mgpbd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  RedWS ws{};
  PcgState* st = nullptr;
  PcgState* st_host = nullptr;   // pinned mirror
  int n_levels = 0, n_smooth = 2, coarse_mode = 0, coarse_sweeps = 50;
  int level_smooth[MGPBD_MAX_LEVELS] = {};   // smoothing steps per phase on level l (0: n_smooth)
  // dense coarsest-level inverse (coarse_mode == MGPBD_COARSE_CHOLESKY)
  double* dense = nullptr;
  size_t dense_cap = 0;
  bool dense_ready = false;
  // Newton-Schulz refresh of the coarsest inverse from the previous one: two steps
  // X <- X + X (I - A X), kept when ||I - A X||_F < 0.1 afterwards, else a full inversion
  // (Newton-Schulz from I / ||A||_inf to convergence)
  bool ns_update = true;
  double* dA = nullptr;      // assembled coarsest matrix
  double* dR = nullptr;      // I - A X
  double* dX2 = nullptr;     // X (Newton-Schulz iterate)
  double* dT = nullptr;      // next X
  double* dnorm = nullptr;   // ||R||_F^2 (device)
  int64_t ns_steps = 0, ns_refactors = 0;
  // the last refresh's acceptance flag (1 accepted, 2 rejected), copied to pinned host memory
  // without a wait and read at the next refresh (ns_event orders it)
  int* ns_flag = nullptr;
  int* ns_flag_host = nullptr;
  hipEvent_t ns_event = nullptr;
  // per-vertex {base tet, vertex weight w'} of the factored level 0's uniform vertex pass
  // (one 8-byte read per vertex; library-owned, grown by mgpbd_mg_set_level0_offsets)
  int2* fac_tbw = nullptr;
  int fac_tbw_cap = 0;
  bool ns_pending = false;
  bool ns_rejected = false;
  float omega = 2.0f / 3.0f;
  int cycle = 1;              // 1: V-cycle, 2: W-cycle (two coarse cycles per level)
  float overcorrect = 1.0f;   // scale of the prolongated coarse correction
  // reduced-precision V-cycle: the dense coarsest product prolongs onto the level above it
  // itself instead of a k_prolong_h launch, and the PCG update writes level 0's scaled rhs
  // and first sweep instead of k_half_first (same values bit for bit)
  bool fused_launch = true;
  // spectral smoother: per-level weights omega_l = target / (safety lambda_l) in device memory,
  // lambda_l from a power iteration on D^-1 A warm-started across updates
  bool spectral = false;
  bool fixed_weights = false;   // per-level weights set by the caller (no spectral refresh)
  bool chebyshev = false;       // spectral weights as a Chebyshev polynomial smoother
  float cheb_ratio = 12.0f;     // smoothed interval [lambda_max / ratio, lambda_max]
  float* cheb_dev = nullptr;    // MGPBD_MAX_LEVELS x kChebSteps x (c1, c2)
  float sm_target = 4.0f / 3.0f, sm_safety = 1.1f;
  int power_iters = 4;
  float* omega_dev = nullptr;   // MGPBD_MAX_LEVELS
  double* pw_dev = nullptr;     // 4 per level: lambda estimate, ||w||_D, previous estimate
  bool pv_ready[MGPBD_MAX_LEVELS] = {};
  int pv_updates[MGPBD_MAX_LEVELS] = {};   // updates since the level was set
  // coarse refill maps, level l -> l + 1 (mgpbd_mg_set_refill), for mgpbd_mg_refresh
  struct Refill {
    int n = -1;              // coarse entries refilled (-1: not registered)
    const int* off = nullptr;
    const int* fpos = nullptr;
    const int* cpos = nullptr;
    const float* fine = nullptr;
    float* coarse = nullptr;
    int n_coarse = 0;
    const int* cdiag = nullptr;
    float* cdinv = nullptr;
  } refill[MGPBD_MAX_LEVELS];
  double* pw_host = nullptr;    // pinned copy of pw_dev from the previous update (read lagged)
  MgLevel L[MGPBD_MAX_LEVELS];
  // factored level 0 of the reduced-precision V-cycle (A^ = K W' K^T + diag(a^), K from the
  // caller's gradients; mgpbd_mg_set_level0_factored): caller-owned arrays
  struct Factored {
    bool on = false;
    int nv = 0;
    const int4* t_i = nullptr;
    const float* invm = nullptr;
    const float* gradC = nullptr;
    const float* alpha = nullptr;
    const int* vs_ptr = nullptr;
    const int* vs_cnt = nullptr;
    const int* vs_inc = nullptr;
    const int* tbase = nullptr;   // per-vertex base tet of the 16-bit slot offsets (or NULL)
    int2* tbw = nullptr;          // {tbase, w'} per vertex, written by k_fac_vert (with tbase)
    int rows_u = 0;               // rows of the uniform vertex-slice layout (0: per slice)
    const uint64_t* tpk = nullptr;  // packed row vertices (or NULL: t_i)
    int pk_bb = 0, pk_ob = 0;       // bits of the first vertex and of each difference
    v2u32* vs_k = nullptr;   // K per vertex slot (4 halves)
    float4* vs_g = nullptr;  // FP32 gradient per vertex slot (the PCG's operator)
    v4u32* tka = nullptr;             // K per tet: halves 0..7 (16-byte plane of n_tets)
    unsigned short* tkz = nullptr;    // K per tet: half 8 (2-byte plane of n_tets)
    float cs = 1.0f, ic2 = 1.0f;      // c = sqrt(max inverse mass) in K, 1 / c^2 in W'
    FacK k() const { return FacK{tka, tkz}; }
    float* ahat = nullptr;   // a^ per tet
    fvec3* u = nullptr;     // per-vertex products of the V-cycle's gather pass
    double4* u64 = nullptr;  // per-vertex products of the PCG operator's gather pass
    float* dinv = nullptr;   // 1 / A_tt from the gradients (level 0's dinv array)
    // level 1 from the gradients (mgpbd_mg_set_level1_factored)
    bool galerkin = false;
    int ng = 0, ne = 0;
    const int* gptr = nullptr;
    const int* gmem = nullptr;
    const int* gvert = nullptr;
    float4* G = nullptr;
    const int* tptr = nullptr;
    const int2* terms = nullptr;
    const int* cpos = nullptr;
    const int* cpos_t = nullptr;
    const int* cagg = nullptr;
    float* cval = nullptr;
  } fac;
  int precision = MGPBD_PREC_FP32;   // V-cycle storage precision
  int flexible = 0;                   // Polak-Ribiere beta
  int poll = 4;
  int verify = 0;   // 1: FP64 check of the converged x, restart from x when it fails; 2: check only
  bool fp64_iterate = false;   // FP64 PCG iterate, FP64-accumulated outer SpMV and restart residual
  // stall rescue (spectral smoothers): a PCG whose residual falls less than 2x over 8
  // iterations re-estimates every level's spectrum from a fresh start vector with the safety
  // factor times rescue_boost and restarts from its iterate (once per solve)
  bool rescue = false;
  float rescue_boost = 1.25f;
  double stall_ratio = 0.25;   // ||r||^2 ratio over 8 iterations counted as a stall
  int64_t n_rescues = 0, n_restarts = 0;
  // lag rescue: levels 1.. may be older than level 0 (mgpbd_mg_refresh with coarse = 0, or
  // mgpbd_mg_set_coarse_lagged); a solve on such a hierarchy that runs past
  // max(8, 2 fresh_iters + 2) iterations rebuilds levels 1.. and resumes from its iterate
  // (once per solve). fresh_iters: the last solve run entirely on a current hierarchy.
  // warm-started solves pick their start by the CG energy instead of the residual
  bool start_energy = false;
  bool coarse_lagged = false;
  int fresh_iters = 0;
  int lag_limit = -1;   // -1: max(8, 2 fresh_iters + 2); 0: no lag rescue; k > 0: k iterations
  int64_t n_lag_rescues = 0;
  // HIP graphs of k PCG iterations (+ the state copy), k = 1 .. kGraphMax, captured when a
  // replay of that length is first needed for the same buffers, replayed on `work` stream
  bool use_graph = true;
  hipStream_t work = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  static constexpr int kGraphMax = 64;
  hipGraphExec_t graphs[kGraphMax + 1] = {};
  const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
  int pred_iters = 0;   // iterations of the previous solve (length of the next first replay)
  // optional per-kernel timing with HIP events on the context stream
  bool prof = false;
  std::vector<hipEvent_t> ev;     // 2 per recorded launch
  std::vector<int> ev_kind;
  int ev_used = 0;
  double prof_ms[MGPBD_PROF_KINDS] = {};
  int64_t prof_cnt[MGPBD_PROF_KINDS] = {};
};

// Every captured iteration graph is dropped (a level, a setting or a buffer changed).
static void drop_graphs(mgpbd_ctx* c) {
  for (hipGraphExec_t& g : c->graphs)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

// Device pointer of level l's smoothing weight, or nullptr for the fixed host weight.
// smoothing steps per phase on level l
static int ns(const mgpbd_ctx* ctx, int l) {
  return ctx->level_smooth[l] > 0 ? ctx->level_smooth[l] : ctx->n_smooth;
}

static const float* wl(const mgpbd_ctx* c, int l) {
  return c->spectral ? c->omega_dev + l : nullptr;
}
// Device pointer of the (c1, c2) pair of Chebyshev step `step` (>= 2) on level l.
static const float* cl(const mgpbd_ctx* c, int l, int step) {
  return c->cheb_dev + 2 * (l * kChebSteps + std::min(step - 2, kChebSteps - 1));
}
// Chebyshev smoothing on level l (not on the coarsest level, which keeps Jacobi sweeps).
static bool cheb_level(const mgpbd_ctx* c, int l) {
  return c->chebyshev && l < c->n_levels - 1 && !c->L[l].diagonal;
}

static int prof_flush(mgpbd_ctx* c) {
  if (c->ev_used == 0) return 0;
  MGPBD_HIP_TRY(hipEventSynchronize(c->ev[2 * (c->ev_used - 1) + 1]));
  for (int i = 0; i < c->ev_used; ++i) {
    float ms = 0.f;
    MGPBD_HIP_TRY(hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
    c->prof_ms[c->ev_kind[i]] += ms;
    c->prof_cnt[c->ev_kind[i]] += 1;
  }
  c->ev_used = 0;
  return 0;
}

// Brackets one launch with events when profiling is on; returns the slot or -1.
static int prof_begin(mgpbd_ctx* c, int kind) {
  if (!c->prof) return -1;
  if (c->ev_used * 2 + 2 > (int)c->ev.size()) {
    if (prof_flush(c)) return -1;
  }
  int i = c->ev_used++;
  c->ev_kind[i] = kind;
  (void)hipEventRecord(c->ev[2 * i], c->stream);
  return i;
}
static void prof_end(mgpbd_ctx* c, int i) {
  if (i >= 0) (void)hipEventRecord(c->ev[2 * i + 1], c->stream);
}

static inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }
static inline int red_blocks(int64_t n) { return std::min(kMaxRedBlocks, nblocks(n)); }

extern "C" {

mgpbd_ctx* mgpbd_ctx_create(void* stream) {
  auto* c = new (std::nothrow) mgpbd_ctx;
  if (!c) { mgpbd_set_error("mgpbd_ctx_create: out of memory"); return nullptr; }
  c->stream = (hipStream_t)stream;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess)
    e = hipMalloc(&c->ws.partials, sizeof(double) * kPartialStride * kPartialSlots);
  if (e == hipSuccess) e = hipMalloc(&c->st, sizeof(PcgState));
  if (e == hipSuccess) e = hipMemset(c->st, 0, sizeof(PcgState));
  if (e == hipSuccess) e = hipHostMalloc(&c->st_host, sizeof(PcgState), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->work, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&c->omega_dev, sizeof(float) * MGPBD_MAX_LEVELS);
  if (e == hipSuccess) e = hipMalloc(&c->pw_dev, sizeof(double) * 4 * MGPBD_MAX_LEVELS);
  if (e == hipSuccess) e = hipMalloc(&c->cheb_dev, sizeof(float) * 2 * kChebSteps * MGPBD_MAX_LEVELS);
  if (e == hipSuccess) e = hipHostMalloc(&c->pw_host, sizeof(double) * 4 * MGPBD_MAX_LEVELS, hipHostMallocDefault);
  if (e == hipSuccess) std::memset(c->pw_host, 0, sizeof(double) * 4 * MGPBD_MAX_LEVELS);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    mgpbd_set_error(std::string("mgpbd_ctx_create: ") + hipGetErrorString(e));
    if (c->ws.partials) (void)hipFree(c->ws.partials);
    if (c->st) (void)hipFree(c->st);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->omega_dev) (void)hipFree(c->omega_dev);
    if (c->pw_dev) (void)hipFree(c->pw_dev);
    if (c->pw_host) (void)hipHostFree(c->pw_host);
    if (c->cheb_dev) (void)hipFree(c->cheb_dev);
    delete c;
    return nullptr;
  }
  return c;
}

int mgpbd_ctx_set_stream(mgpbd_ctx* c, void* stream) {
  if (!c) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_set_stream: null context");
  c->stream = (hipStream_t)stream;
  return 0;
}

void mgpbd_ctx_destroy(mgpbd_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->work) (void)hipStreamDestroy(c->work);
  if (c->dense) (void)hipFree(c->dense);
  if (c->dA) (void)hipFree(c->dA);
  if (c->dR) (void)hipFree(c->dR);
  if (c->dX2) (void)hipFree(c->dX2);
  if (c->dT) (void)hipFree(c->dT);
  if (c->dnorm) (void)hipFree(c->dnorm);
  if (c->ns_flag) (void)hipFree(c->ns_flag);
  if (c->fac_tbw) (void)hipFree(c->fac_tbw);
  if (c->ns_flag_host) (void)hipHostFree(c->ns_flag_host);
  if (c->ns_event) (void)hipEventDestroy(c->ns_event);
  (void)hipFree(c->ws.partials);
  (void)hipFree(c->st);
  (void)hipHostFree(c->st_host);
  (void)hipFree(c->omega_dev);
  (void)hipFree(c->pw_dev);
  (void)hipHostFree(c->pw_host);
  (void)hipFree(c->cheb_dev);
  delete c;
}

int mgpbd_ctx_device_bytes(mgpbd_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_device_bytes: invalid arguments");
  int64_t t = (int64_t)sizeof(double) * kPartialStride * kPartialSlots +
              (int64_t)sizeof(PcgState) +
              (int64_t)sizeof(float) * MGPBD_MAX_LEVELS * (1 + 2 * kChebSteps) +
              (int64_t)sizeof(double) * 4 * MGPBD_MAX_LEVELS;
  int nd = 0;
  for (double* b : {c->dense, c->dA, c->dR, c->dX2, c->dT}) nd += b != nullptr;
  t += (int64_t)nd * (int64_t)c->dense_cap * (int64_t)sizeof(double);
  t += (int64_t)c->fac_tbw_cap * (int64_t)sizeof(int2);
  *bytes = t;
  return 0;
}

int mgpbd_ctx_sync(mgpbd_ctx* c) {
  if (!c) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_sync: null context");
  MGPBD_HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int mgpbd_ctx_profile(mgpbd_ctx* c, int enable) {
  if (!c) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_profile: null context");
  if (int e = prof_flush(c)) return e;
  if (enable && c->ev.empty()) {
    c->ev.resize(2 * 2048);
    c->ev_kind.resize(2048);
    for (auto& e : c->ev) MGPBD_HIP_TRY(hipEventCreate(&e));
  }
  c->prof = enable != 0;
  for (int k = 0; k < MGPBD_PROF_KINDS; ++k) { c->prof_ms[k] = 0.0; c->prof_cnt[k] = 0; }
  return 0;
}

int mgpbd_ctx_profile_read(mgpbd_ctx* c, double* ms, int64_t* count) {
  if (!c || !ms || !count) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_profile_read: invalid arguments");
  MGPBD_HIP_TRY(hipStreamSynchronize(c->stream));
  if (int e = prof_flush(c)) return e;
  for (int k = 0; k < MGPBD_PROF_KINDS; ++k) { ms[k] = c->prof_ms[k]; count[k] = c->prof_cnt[k]; }
  return 0;
}

#define CTX_CHECK(fn)                                                           \
  if (!ctx) return mgpbd_fail(MGPBD_EINVAL, std::string(fn) + ": null context")

int mgpbd_rest_shape(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_ref,
                     const float* mu_t, double mu, double dt, float* Dm_inv, float* vol,
                     float* alpha) {
  CTX_CHECK("mgpbd_rest_shape");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_rest_shape: n_tets < 0");
  if (n_tets == 0) return 0;
  k_rest_shape<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_ref, mu_t, mu, nullptr, 0.0, 0, dt, Dm_inv,
      vol, alpha);
  MGPBD_LAUNCH_CHECK("mgpbd_rest_shape");
  return 0;
}

int mgpbd_rest_shape_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_ref,
                         const float* mu_t, double mu, const float* lam_t, double lam, double dt,
                         float* Dm_inv, float* vol, float* alpha) {
  CTX_CHECK("mgpbd_rest_shape_vol");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_rest_shape_vol: n_tets < 0");
  if (!lam_t && !(lam > 0.0))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_rest_shape_vol: lambda must be > 0 (or per tet)");
  if (n_tets == 0) return 0;
  k_rest_shape<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_ref, mu_t, mu, lam_t, lam, 1, dt, Dm_inv,
      vol, alpha);
  MGPBD_LAUNCH_CHECK("mgpbd_rest_shape_vol");
  return 0;
}

int mgpbd_predict(mgpbd_ctx* ctx, int n_verts, float* v_p, float* v_v, float* v_p_old,
                  const float* invm, const float* fext, float gx, float gy, float gz, float dt,
                  float damp, int hold_pinned) {
  CTX_CHECK("mgpbd_predict");
  if (n_verts <= 0) return n_verts < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_predict: n < 0") : 0;
  k_predict<<<nblocks(n_verts), kBlock, 0, ctx->stream>>>(n_verts, v_p, v_v, v_p_old, invm,
                                                          fext, gx, gy, gz, dt, damp, hold_pinned);
  MGPBD_LAUNCH_CHECK("mgpbd_predict");
  return 0;
}

int mgpbd_spring_drag(mgpbd_ctx* ctx, int n_verts, const float* v_p, float* f_ext, int vertex,
                      float tx, float ty, float tz, float k) {
  CTX_CHECK("mgpbd_spring_drag");
  if (n_verts <= 0 || !v_p || !f_ext || vertex < 0 || vertex >= n_verts)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_spring_drag: invalid arguments");
  k_spring_drag<<<nblocks(3LL * n_verts), kBlock, 0, ctx->stream>>>(n_verts, v_p, f_ext, vertex,
                                                                    tx, ty, tz, k);
  MGPBD_LAUNCH_CHECK("mgpbd_spring_drag");
  return 0;
}

int mgpbd_update_vel(mgpbd_ctx* ctx, int n_verts, const float* v_p, const float* v_p_old,
                     float* v_v, float dt, float damp) {
  CTX_CHECK("mgpbd_update_vel");
  if (n_verts <= 0) return n_verts < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_update_vel: n < 0") : 0;
  k_update_vel<<<nblocks(3LL * n_verts), kBlock, 0, ctx->stream>>>(n_verts, v_p, v_p_old, v_v, dt, damp);
  MGPBD_LAUNCH_CHECK("mgpbd_update_vel");
  return 0;
}

int mgpbd_constraints(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                      const float* Dm_inv, const float* alpha, const float* lambda, float* C,
                      float* gradC, float* rhs, double* norm2) {
  CTX_CHECK("mgpbd_constraints");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints: n_tets < 0");
  int pe = prof_begin(ctx, MGPBD_PROF_CONSTRAINTS);
  const int g = red_blocks(n_tets);
  k_constraints<false><<<g, kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_p, Dm_inv, alpha, lambda, C, gradC, rhs,
      ctx->ws);
  k_fin_sum<<<1, kBlock, 0, ctx->stream>>>(ctx->ws, g, 0, 1, norm2);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_constraints");
  return 0;
}

int mgpbd_constraints_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                          const float* Dm_inv, const float* alpha, const float* lambda, float* C,
                          float* gradC, float* rhs, double* norm2) {
  CTX_CHECK("mgpbd_constraints_vol");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints_vol: n_tets < 0");
  int pe = prof_begin(ctx, MGPBD_PROF_CONSTRAINTS);
  const int g = red_blocks(n_tets);
  k_constraints<true><<<g, kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_p, Dm_inv, alpha, lambda, C, gradC, rhs,
      ctx->ws);
  k_fin_sum<<<1, kBlock, 0, ctx->stream>>>(ctx->ws, g, 0, 1, norm2);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_constraints_vol");
  return 0;
}

int mgpbd_fill_A(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr, const int32_t* idx,
                 const int32_t* row_len, const int32_t* t_i, const float* invm,
                 const float* gradC, const float* alpha, float* val, float* dinv) {
  CTX_CHECK("mgpbd_fill_A");
  if (n_tets <= 0) return n_tets < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_fill_A: n < 0") : 0;
  int pe = prof_begin(ctx, MGPBD_PROF_FILL_A);
  k_fill_A<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(
      n_tets, slice_ptr, idx, row_len, reinterpret_cast<const int4*>(t_i), invm, gradC, alpha,
      val, dinv);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_fill_A");
  return 0;
}

int mgpbd_fill_A_desc(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr, const int32_t* idx,
                      const int32_t* row_len, const int32_t* t_i, uint16_t* desc) {
  CTX_CHECK("mgpbd_fill_A_desc");
  if (n_tets < 0 || (n_tets > 0 && (!slice_ptr || !idx || !row_len || !t_i || !desc)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_fill_A_desc: invalid arguments");
  if (n_tets == 0) return 0;
  k_fill_desc<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(
      n_tets, slice_ptr, idx, row_len, reinterpret_cast<const int4*>(t_i), desc);
  MGPBD_LAUNCH_CHECK("mgpbd_fill_A_desc");
  return 0;
}

int mgpbd_fill_A_shared(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr,
                        const int32_t* idx, const int32_t* row_len, const uint16_t* desc,
                        const int32_t* t_i, const float* invm, const float* gradC,
                        const float* alpha, float* val, float* dinv) {
  CTX_CHECK("mgpbd_fill_A_shared");
  if (n_tets <= 0) return n_tets < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_fill_A_shared: n < 0") : 0;
  if (!desc) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_fill_A_shared: no descriptors");
  int pe = prof_begin(ctx, MGPBD_PROF_FILL_A);
  k_fill_A_desc<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(
      n_tets, slice_ptr, idx, row_len, desc, reinterpret_cast<const int4*>(t_i), invm, gradC,
      alpha, val, dinv);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_fill_A_shared");
  return 0;
}

int mgpbd_scatter(mgpbd_ctx* ctx, int n_tets, int n_verts, const int32_t* vptr,
                  const int32_t* vinc, const float* invm, const float* gradC, const float* dlam,
                  float* lambda, float* dpos) {
  CTX_CHECK("mgpbd_scatter");
  if (n_tets < 0 || n_verts < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_scatter: n < 0");
  if (n_tets > 0) {
    k_add_inplace<<<nblocks(n_tets), kBlock, 0, ctx->stream>>>(n_tets, dlam, lambda);
    MGPBD_LAUNCH_CHECK("mgpbd_scatter(lambda)");
  }
  if (n_verts > 0) {
    k_scatter<<<nblocks(n_verts), kBlock, 0, ctx->stream>>>(n_verts, vptr, vinc, invm, gradC, dlam, dpos);
    MGPBD_LAUNCH_CHECK("mgpbd_scatter");
  }
  return 0;
}

int mgpbd_dual_norms(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                     const float* dpos, float omega, const float* Dm_inv, const float* alpha,
                     const float* lambda, const float* C, int with_current, double* out) {
  CTX_CHECK("mgpbd_dual_norms");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_dual_norms: n < 0");
  const int g = red_blocks(n_tets);
  k_dual_norms<false><<<g, kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_p, dpos, omega, Dm_inv, alpha, lambda, C,
      with_current, ctx->ws);
  k_fin_sum<<<1, kBlock, 0, ctx->stream>>>(ctx->ws, g, with_current ? 0 : 1, 2, out);
  MGPBD_LAUNCH_CHECK("mgpbd_dual_norms");
  return 0;
}

int mgpbd_dual_norms_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                         const float* dpos, float omega, const float* Dm_inv, const float* alpha,
                         const float* lambda, const float* C, int with_current, double* out) {
  CTX_CHECK("mgpbd_dual_norms_vol");
  if (n_tets < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_dual_norms_vol: n < 0");
  const int g = red_blocks(n_tets);
  k_dual_norms<true><<<g, kBlock, 0, ctx->stream>>>(
      n_tets, reinterpret_cast<const int4*>(t_i), v_p, dpos, omega, Dm_inv, alpha, lambda, C,
      with_current, ctx->ws);
  k_fin_sum<<<1, kBlock, 0, ctx->stream>>>(ctx->ws, g, with_current ? 0 : 1, 2, out);
  MGPBD_LAUNCH_CHECK("mgpbd_dual_norms_vol");
  return 0;
}

int mgpbd_apply_dpos(mgpbd_ctx* ctx, int n_verts, float* v_p, const float* dpos,
                     const float* invm, float omega) {
  CTX_CHECK("mgpbd_apply_dpos");
  if (n_verts <= 0) return n_verts < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_apply_dpos: n < 0") : 0;
  k_apply_dpos<<<nblocks(n_verts), kBlock, 0, ctx->stream>>>(n_verts, v_p, dpos, invm, omega);
  MGPBD_LAUNCH_CHECK("mgpbd_apply_dpos");
  return 0;
}

int mgpbd_spmv(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr, const int32_t* idx,
               const float* val, const float* x, float* y) {
  CTX_CHECK("mgpbd_spmv");
  if (n <= 0) return n < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_spmv: n < 0") : 0;
  int ns = (n + 63) / 64;
  k_sell<0><<<nblocks(64LL * ns), kBlock, 0, ctx->stream>>>(n, ns, slice_ptr, idx, val, x, y,
                                                           nullptr, nullptr, 0.f, nullptr, nullptr);
  MGPBD_LAUNCH_CHECK("mgpbd_spmv");
  return 0;
}

int mgpbd_jacobi(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr, const int32_t* idx,
                 const float* val, const float* dinv, const float* b, const float* x,
                 float* x_new, float omega) {
  CTX_CHECK("mgpbd_jacobi");
  if (n <= 0) return n < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_jacobi: n < 0") : 0;
  if (x == x_new) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_jacobi: x and x_new must differ");
  int ns = (n + 63) / 64;
  k_sell<1><<<nblocks(64LL * ns), kBlock, 0, ctx->stream>>>(n, ns, slice_ptr, idx, val, x, x_new,
                                                           dinv, b, omega, nullptr, nullptr);
  MGPBD_LAUNCH_CHECK("mgpbd_jacobi");
  return 0;
}

int mgpbd_diag_inv(mgpbd_ctx* ctx, int n, const int32_t* diag_pos, const float* val, float* dinv) {
  CTX_CHECK("mgpbd_diag_inv");
  if (n <= 0) return n < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_diag_inv: n < 0") : 0;
  k_diag_inv<<<nblocks(n), kBlock, 0, ctx->stream>>>(n, diag_pos, val, dinv);
  MGPBD_LAUNCH_CHECK("mgpbd_diag_inv");
  return 0;
}

int mgpbd_coarse_refill(mgpbd_ctx* ctx, int nc, const int32_t* off, const int32_t* fpos,
                        const int32_t* cpos, const float* fine, float* coarse) {
  CTX_CHECK("mgpbd_coarse_refill");
  if (nc <= 0) return nc < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_coarse_refill: n < 0") : 0;
  int pe = prof_begin(ctx, MGPBD_PROF_REFILL);
  k_refill<<<nblocks(nc), kBlock, 0, ctx->stream>>>(nc, off, fpos, cpos, fine, coarse);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_coarse_refill");
  return 0;
}

int mgpbd_restrict(mgpbd_ctx* ctx, int nc, const int32_t* mptr, const int32_t* mem,
                   const float* r, float* bc) {
  CTX_CHECK("mgpbd_restrict");
  if (nc <= 0) return nc < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_restrict: n < 0") : 0;
  k_restrict<<<nblocks(nc), kBlock, 0, ctx->stream>>>(nc, mptr, mem, r, bc, nullptr, nullptr, 0.f, nullptr, nullptr);
  MGPBD_LAUNCH_CHECK("mgpbd_restrict");
  return 0;
}

int mgpbd_prolong(mgpbd_ctx* ctx, int n, const int32_t* agg, const float* xc, float* x) {
  CTX_CHECK("mgpbd_prolong");
  if (n <= 0) return n < 0 ? mgpbd_fail(MGPBD_EINVAL, "mgpbd_prolong: n < 0") : 0;
  k_prolong<<<nblocks(n), kBlock, 0, ctx->stream>>>(n, agg, xc, x, 1.0f, nullptr);
  MGPBD_LAUNCH_CHECK("mgpbd_prolong");
  return 0;
}

static int blas1(mgpbd_ctx* ctx, const char* name, int op, int n, float alpha, const float* a,
                 float* y) {
  CTX_CHECK(name);
  if (n <= 0) return n < 0 ? mgpbd_fail(MGPBD_EINVAL, std::string(name) + ": n < 0") : 0;
  k_blas1<<<nblocks(n), kBlock, 0, ctx->stream>>>(op, n, alpha, a, y);
  MGPBD_LAUNCH_CHECK(name);
  return 0;
}

int mgpbd_axpy(mgpbd_ctx* ctx, int n, float alpha, const float* x, float* y) { return blas1(ctx, "mgpbd_axpy", 0, n, alpha, x, y); }
int mgpbd_xpay(mgpbd_ctx* ctx, int n, float alpha, const float* x, float* y) { return blas1(ctx, "mgpbd_xpay", 1, n, alpha, x, y); }
int mgpbd_copy(mgpbd_ctx* ctx, int n, const float* src, float* dst) { return blas1(ctx, "mgpbd_copy", 2, n, 0.f, src, dst); }
int mgpbd_fill(mgpbd_ctx* ctx, int n, float value, float* x) { return blas1(ctx, "mgpbd_fill", 3, n, value, nullptr, x); }
int mgpbd_scale(mgpbd_ctx* ctx, int n, float alpha, float* x) { return blas1(ctx, "mgpbd_scale", 4, n, alpha, nullptr, x); }
int mgpbd_negate_and_add(mgpbd_ctx* ctx, int n, const float* b, float* r) { return blas1(ctx, "mgpbd_negate_and_add", 5, n, 0.f, b, r); }

int mgpbd_dot(mgpbd_ctx* ctx, int n, const float* a, const float* b, double* out) {
  CTX_CHECK("mgpbd_dot");
  if (n < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_dot: n < 0");
  const int g = red_blocks(n);
  k_dot<<<g, kBlock, 0, ctx->stream>>>(n, a, b, ctx->ws);
  k_fin_sum<<<1, kBlock, 0, ctx->stream>>>(ctx->ws, g, 0, 1, out);
  MGPBD_LAUNCH_CHECK("mgpbd_dot");
  return 0;
}

// ============================================================================================
// Multigrid V-cycle and PCG
// ============================================================================================
int mgpbd_mg_configure(mgpbd_ctx* ctx, int n_levels, int n_smooth, float omega, int coarse_mode,
                       int coarse_sweeps) {
  CTX_CHECK("mgpbd_mg_configure");
  if (n_levels < 1 || n_levels > MGPBD_MAX_LEVELS || n_smooth < 0 || coarse_sweeps < 1 ||
      (coarse_mode != MGPBD_COARSE_JACOBI && coarse_mode != MGPBD_COARSE_CHOLESKY))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_configure: invalid arguments");
  ctx->dense_ready = false;
  drop_graphs(ctx);
  ctx->n_levels = n_levels;
  ctx->n_smooth = n_smooth;
  for (int l = 0; l < MGPBD_MAX_LEVELS; ++l) ctx->level_smooth[l] = 0;
  ctx->omega = omega;
  ctx->coarse_mode = coarse_mode;
  ctx->coarse_sweeps = coarse_sweeps;
  for (int l = 0; l < MGPBD_MAX_LEVELS; ++l) {
    ctx->L[l] = MgLevel{};
    ctx->pv_ready[l] = false;
  }
  return 0;
}

int mgpbd_mg_set_level(mgpbd_ctx* ctx, int level, int n_rows, const int32_t* slice_ptr,
                       const int32_t* idx, const float* val, const float* dinv,
                       const int32_t* agg, const int32_t* mptr, const int32_t* mem,
                       float* work, int diagonal_only) {
  CTX_CHECK("mgpbd_mg_set_level");
  if (level < 0 || level >= ctx->n_levels || n_rows <= 0 || !slice_ptr || !idx || !val ||
      !dinv || !work)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level: invalid arguments");
  bool coarsest = (level == ctx->n_levels - 1);
  if (!coarsest && (!agg || !mptr || !mem))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level: transfer arrays required above the coarsest level");
  if (coarsest && ctx->coarse_mode == MGPBD_COARSE_CHOLESKY && n_rows > MGPBD_MAX_DENSE)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level: coarsest level too large for the dense "
                                    "solve (" + std::to_string(n_rows) + " > " +
                                    std::to_string(MGPBD_MAX_DENSE) + ")");
  drop_graphs(ctx);
  MgLevel& L = ctx->L[level];
  L.n = n_rows;
  L.n_slices = (n_rows + 63) / 64;
  L.slice_ptr = slice_ptr; L.idx = idx; L.val = val; L.dinv = dinv;
  L.agg = agg; L.mptr = mptr; L.mem = mem; L.work = work;
  L.diagonal = diagonal_only;
  L.seg_ptr = nullptr; L.segs = nullptr; L.win_segs = 0; L.woff = nullptr;
  if (level == 0) ctx->fac = mgpbd_ctx::Factored{};
  return 0;
}

static int check_levels(mgpbd_ctx* ctx) {
  if (ctx->n_levels < 1) return mgpbd_fail(MGPBD_EINVAL, "multigrid: not configured");
  for (int l = 0; l < ctx->n_levels; ++l)
    if (ctx->L[l].n <= 0) return mgpbd_fail(MGPBD_EINVAL, "multigrid: level " + std::to_string(l) + " not set");
  return 0;
}

static int check_ready(mgpbd_ctx* ctx) {
  if (int e = check_levels(ctx)) return e;
  if (ctx->coarse_mode == MGPBD_COARSE_CHOLESKY && !ctx->dense_ready)
    return mgpbd_fail(MGPBD_EINVAL, "multigrid: coarsest level not factored "
                                    "(call mgpbd_mg_factor_coarsest after its values change)");
  if (ctx->coarse_mode == MGPBD_COARSE_CHOLESKY && ctx->n_levels < 2)
    return mgpbd_fail(MGPBD_EINVAL, "multigrid: the dense coarsest solve needs two levels or more");
  if (ctx->precision != MGPBD_PREC_FP32)
    for (int l = 0; l < ctx->n_levels; ++l)
      if (!ctx->L[l].hval && !ctx->L[l].hpk && !(l == 0 && ctx->fac.on))
        return mgpbd_fail(MGPBD_EINVAL, "multigrid: reduced precision without half buffers on level " + std::to_string(l));
  return 0;
}

// FP32 row kernels of a level: windowed when the level has FP32 LDS offsets
extern "C++" {
template <int MODE>
static void sell_launch(mgpbd_ctx* ctx, const MgLevel& L, const float* x, float* y,
                        const float* dinv, const float* b, float omega, const float* wp,
                        const int* done, const float* xp = nullptr, const float* cc = nullptr) {
  const int g = nblocks(64LL * L.n_slices);
  if (L.woff)
    k_sellw_f32<MODE><<<g, kBlock, sizeof(float) * kWinSeg * L.win_segs, ctx->stream>>>(
        L.n, L.n_slices, L.slice_ptr, L.woff, L.val, L.seg_ptr, L.segs, x, y, dinv, b, omega, wp,
        done, xp, cc);
  else
    k_sell<MODE><<<g, kBlock, 0, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.val, x, y,
                                               dinv, b, omega, wp, done, xp, cc);
}
}  // extern "C++"

// Jacobi sweeps on level l from the first-sweep result already in *cur; returns the buffer
// holding the result (cur or the level's other x buffer).
// `count` smoothing steps on level l from the iterate in `cur` (one of the level's two x
// buffers). first_step = the 1-based index of the first step within its smoothing phase (a
// pre-smoothing phase from zero enters at step 2: its first step is fused elsewhere), 0 for
// plain weighted-Jacobi sweeps (the coarsest level). With the Chebyshev smoother steps >= 2
// need the previous iterate, which is the buffer being overwritten (zero at step 2 of a phase
// that started from zero).
static float* sweeps(mgpbd_ctx* ctx, int l, const float* b, float* cur, int count, const int* done,
                     int first_step = 0, bool from_zero = false) {
  MgLevel& L = ctx->L[l];
  float* xa = L.wb(1);
  float* xb = L.wb(2);
  if (count <= 0) return cur;
  if (L.diagonal) {
    k_jacobi_diag<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.slice_ptr, L.val, L.dinv, b,
                                                            cur, ctx->omega, wl(ctx, l), count, done);
    return cur;
  }
  if (L.n <= kSmallRows && count >= 8) {   // many sweeps of a small level: one launch
    int pe = prof_begin(ctx, MGPBD_PROF_COARSE);
    k_jacobi_small<<<1, kSmallBlock, 0, ctx->stream>>>(L.n, L.slice_ptr, L.idx, L.val, L.dinv,
                                                       b, cur, ctx->omega, wl(ctx, l), count, done);
    prof_end(ctx, pe);
    return cur;
  }
  const bool cheb = first_step > 0 && cheb_level(ctx, l);
  for (int s = 0; s < count; ++s) {
    float* nxt = (cur == xa) ? xb : xa;
    const int step = first_step + s;
    int pe = prof_begin(ctx, l == 0 ? MGPBD_PROF_JACOBI_L0 : MGPBD_PROF_COARSE);
    if (cheb && step >= 2) {
      const float* xp = (step == 2 && from_zero) ? nullptr : nxt;
      sell_launch<3>(ctx, L, cur, nxt, L.dinv, b, ctx->omega, nullptr, done, xp, cl(ctx, l, step));
    } else {
      sell_launch<1>(ctx, L, cur, nxt, L.dinv, b, ctx->omega, wl(ctx, l), done);
    }
    prof_end(ctx, pe);
    cur = nxt;
  }
  return cur;
}

// Counts the outcome of the last Newton-Schulz refresh once its flag has arrived (the event
// completed long before in the solve's own synchronisations).
static int ns_settle(mgpbd_ctx* ctx) {
  if (!ctx->ns_pending) return 0;
  MGPBD_HIP_TRY(hipEventSynchronize(ctx->ns_event));
  ctx->ns_pending = false;
  if (*ctx->ns_flag_host == 1) ctx->ns_steps += 2;
  else ctx->ns_rejected = true;
  return 0;
}

static int convert_half_levels(mgpbd_ctx* ctx, int first, int last);

int mgpbd_mg_set_refill(mgpbd_ctx* ctx, int level, int n_entries, const int32_t* off,
                        const int32_t* fpos, const int32_t* cpos, const float* fine_val,
                        float* coarse_val, int n_coarse, const int32_t* coarse_diag_pos,
                        float* coarse_dinv) {
  CTX_CHECK("mgpbd_mg_set_refill");
  if (level < 0 || level >= MGPBD_MAX_LEVELS - 1 || n_entries < 0 || n_coarse < 0 ||
      (n_entries > 0 && (!off || !fpos || !cpos || !fine_val || !coarse_val)) ||
      (n_coarse > 0 && (!coarse_diag_pos || !coarse_dinv || !coarse_val)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_refill: invalid arguments");
  mgpbd_ctx::Refill& R = ctx->refill[level];
  R.n = n_entries; R.off = off; R.fpos = fpos; R.cpos = cpos; R.fine = fine_val;
  R.coarse = coarse_val; R.n_coarse = n_coarse; R.cdiag = coarse_diag_pos; R.cdinv = coarse_dinv;
  return 0;
}

// Levels 1.. from level 0's current values through the registered refill maps (1/diag of
// each), the coarsest inverse refreshed when `factor`, and the reduced-precision copies of
// levels half_from.. converted.
static int rebuild_coarse(mgpbd_ctx* ctx, bool factor, int half_from) {
  for (int l = 0; l + 1 < ctx->n_levels; ++l) {
    const mgpbd_ctx::Refill& R = ctx->refill[l];
    if (R.n < 0) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_refresh: refill maps of level " +
                                                     std::to_string(l) + " not registered");
    if (l == 0 && ctx->fac.on) {
      if (int e = mgpbd_mg_factored_refill(ctx)) return e;
    } else if (R.n > 0) {
      if (int e = mgpbd_coarse_refill(ctx, R.n, R.off, R.fpos, R.cpos, R.fine, R.coarse))
        return e;
    }
    if (int e = mgpbd_diag_inv(ctx, R.n_coarse, R.cdiag, R.coarse, R.cdinv)) return e;
  }
  if (factor && ctx->coarse_mode == MGPBD_COARSE_CHOLESKY && ctx->n_levels > 1)
    if (int e = mgpbd_mg_factor_coarsest(ctx)) return e;
  if (int e = convert_half_levels(ctx, half_from, ctx->n_levels)) return e;
  ctx->coarse_lagged = false;
  return 0;
}

int mgpbd_mg_refresh(mgpbd_ctx* ctx, int refresh, int half_from, int coarse) {
  CTX_CHECK("mgpbd_mg_refresh");
  if (int e = check_levels(ctx)) return e;
  if (half_from < 0 || half_from > ctx->n_levels)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_refresh: 0 <= half_from <= n_levels");
  if (!coarse) {   // levels 1.. kept: only level 0's own reduced-precision data
    if (half_from == 0)
      if (int e = convert_half_levels(ctx, 0, 1)) return e;
    ctx->coarse_lagged = ctx->n_levels > 1;
    if (refresh && ctx->spectral && !ctx->fixed_weights)
      if (int e = mgpbd_mg_update_smoother(ctx, nullptr)) return e;
    return 0;
  }
  if (int e = rebuild_coarse(ctx, refresh != 0, half_from)) return e;
  if (refresh && ctx->spectral && !ctx->fixed_weights)
    if (int e = mgpbd_mg_update_smoother(ctx, nullptr)) return e;
  return 0;
}

int mgpbd_mg_set_coarse_lagged(mgpbd_ctx* ctx, int lagged) {
  CTX_CHECK("mgpbd_mg_set_coarse_lagged");
  ctx->coarse_lagged = lagged != 0 && ctx->n_levels > 1;
  return 0;
}

// One Newton-Schulz step X <- X + X (I - A X) (T receives the new X; R = I - A X on exit).
static void ns_step(mgpbd_ctx* ctx, int n, const double* X, double* T) {
  hipStream_t s = ctx->stream;
  const size_t need = (size_t)n * n;
  const dim3 gg((n + kGemmTile - 1) / kGemmTile, (n + kGemmTile - 1) / kGemmTile);
  k_set_identity<<<nblocks((int64_t)need), kBlock, 0, s>>>(n, ctx->dR);
  k_dgemm<<<gg, kBlock, 0, s>>>(n, -1.0, ctx->dA, X, 1.0, ctx->dR);
  (void)hipMemcpyAsync(T, X, need * sizeof(double), hipMemcpyDeviceToDevice, s);
  k_dgemm<<<gg, kBlock, 0, s>>>(n, 1.0, X, ctx->dR, 1.0, T);
}

// ||I - A X||_F^2 into ctx->dnorm (queued)
static void ns_residual(mgpbd_ctx* ctx, int n, const double* X) {
  hipStream_t s = ctx->stream;
  const size_t need = (size_t)n * n;
  const dim3 gg((n + kGemmTile - 1) / kGemmTile, (n + kGemmTile - 1) / kGemmTile);
  const int gq = std::min(nblocks((int64_t)need), 1024);
  k_set_identity<<<nblocks((int64_t)need), kBlock, 0, s>>>(n, ctx->dR);
  k_dgemm<<<gg, kBlock, 0, s>>>(n, -1.0, ctx->dA, X, 1.0, ctx->dR);
  k_sumsq<<<gq, kBlock, 0, s>>>((int64_t)need, ctx->dR, ctx->ws);
  k_fin_sum<<<1, kBlock, 0, s>>>(ctx->ws, gq, 0, 1, ctx->dnorm);
}

int mgpbd_mg_factor_coarsest(mgpbd_ctx* ctx) {
  CTX_CHECK("mgpbd_mg_factor_coarsest");
  if (int e = check_levels(ctx)) return e;
  if (ctx->coarse_mode != MGPBD_COARSE_CHOLESKY) return 0;
  MgLevel& L = ctx->L[ctx->n_levels - 1];
  const int n = L.n;
  const size_t need = (size_t)n * n;
  hipStream_t s = ctx->stream;
  if (need > ctx->dense_cap) {
    for (double** b : {&ctx->dense, &ctx->dA, &ctx->dR, &ctx->dX2, &ctx->dT}) {
      if (*b) MGPBD_HIP_TRY(hipFree(*b));
      *b = nullptr;
    }
    for (double** b : {&ctx->dense, &ctx->dA, &ctx->dR, &ctx->dX2, &ctx->dT})
      MGPBD_HIP_TRY(hipMalloc(b, need * sizeof(double)));
    ctx->dense_cap = need;
    ctx->dense_ready = false;
  }
  if (!ctx->dnorm) MGPBD_HIP_TRY(hipMalloc(&ctx->dnorm, sizeof(double)));
  k_dense_zero<<<nblocks((int64_t)n * n), kBlock, 0, s>>>((int64_t)n * n, ctx->dA);
  k_dense_from_sell<<<(n + 3) / 4, kBlock, 0, s>>>(n, L.slice_ptr, L.idx, L.val, ctx->dA);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_factor_coarsest(assemble)");
  bool refactor = !ctx->dense_ready || !ctx->ns_update;
  if (int e = ns_settle(ctx)) return e;
  if (ctx->ns_rejected) {   // the previous refresh's solves ran on the older inverse
    ctx->ns_rejected = false;
    refactor = true;
  }
  if (!refactor) {
    // two Newton-Schulz steps from the previous inverse, then an a posteriori check:
    // ||I - A X||_F < 0.1 bounds the spectral radius of I - A X, so the kept X is SPD with
    // every eigenvalue of X A within 0.1 of 1
    double* X = ctx->dX2;
    double* T = ctx->dT;
    MGPBD_HIP_TRY(hipMemcpyAsync(X, ctx->dense, need * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int step = 0; step < 2; ++step) {
      ns_step(ctx, n, X, T);
      std::swap(X, T);
    }
    ns_residual(ctx, n, X);
    // accepted X is copied (not swapped: captured PCG graphs hold the address of ctx->dense)
    // on the device, without a host wait; a rejected one leaves the previous inverse (SPD, so
    // the V-cycle stays SPD; the verified PCG still reaches its tolerance) and the next
    // refresh inverts from scratch
    if (!ctx->ns_flag) MGPBD_HIP_TRY(hipMalloc(&ctx->ns_flag, sizeof(int)));
    if (!ctx->ns_flag_host) MGPBD_HIP_TRY(hipHostMalloc(&ctx->ns_flag_host, sizeof(int)));
    if (!ctx->ns_event) MGPBD_HIP_TRY(hipEventCreateWithFlags(&ctx->ns_event, hipEventDisableTiming));
    k_ns_accept<<<nblocks((int64_t)need), kBlock, 0, s>>>((int64_t)need, ctx->dnorm, X,
                                                          ctx->dense, ctx->ns_flag);
    MGPBD_LAUNCH_CHECK("mgpbd_mg_factor_coarsest(newton-schulz)");
    MGPBD_HIP_TRY(hipMemcpyAsync(ctx->ns_flag_host, ctx->ns_flag, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    MGPBD_HIP_TRY(hipEventRecord(ctx->ns_event, s));
    ctx->ns_pending = true;
  }
  if (refactor) {
    // Full inversion: Newton-Schulz from X0 = I / ||A||_inf. For SPD A every eigenvalue of
    // X0 A lies in (0, 1], so ||I - X A||_2 < 1 and the iteration converges (quadratically
    // once below 1). Rounds of four steps, each followed by a host check of ||I - A X||_F:
    // done when it falls below 1e-10 or stops falling below 0.1 (rounding floor); an
    // indefinite or singular level never converges (MGPBD_ENONFINITE after 64 steps).
    ++ctx->ns_refactors;
    double* X = ctx->dX2;
    double* T = ctx->dT;
    k_inf_norm<<<1, kBlock, 0, s>>>(n, ctx->dA, ctx->dnorm);
    k_set_scaled_identity<<<nblocks((int64_t)need), kBlock, 0, s>>>(n, ctx->dnorm, X);
    double prev = INFINITY, f = INFINITY;
    bool done = false;
    for (int round = 0; round < 16 && !done; ++round) {
      for (int step = 0; step < 4; ++step) {
        ns_step(ctx, n, X, T);
        std::swap(X, T);
      }
      ns_residual(ctx, n, X);
      MGPBD_LAUNCH_CHECK("mgpbd_mg_factor_coarsest(inversion)");
      double f2 = 0.0;
      MGPBD_HIP_TRY(hipMemcpyAsync(&f2, ctx->dnorm, sizeof(double), hipMemcpyDeviceToHost, s));
      MGPBD_HIP_TRY(hipStreamSynchronize(s));
      f = std::sqrt(f2);
      if (!std::isfinite(f) || f > 1e6) break;
      done = f < 1e-10 || (f < 0.1 && f > 0.5 * prev);
      prev = f;
    }
    if (!done) {
      ctx->dense_ready = false;
      return mgpbd_fail(MGPBD_ENONFINITE, "coarsest level: the Newton-Schulz inverse did not "
                                          "converge (not positive definite?), ||I - A X||_F = " +
                                              std::to_string(f));
    }
    MGPBD_HIP_TRY(hipMemcpyAsync(ctx->dense, X, need * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  ctx->dense_ready = true;
  return 0;
}

int mgpbd_mg_coarse_stats(mgpbd_ctx* ctx, int64_t* newton_schulz_steps, int64_t* refactors) {
  CTX_CHECK("mgpbd_mg_coarse_stats");
  if (int e = ns_settle(ctx)) return e;
  if (newton_schulz_steps) *newton_schulz_steps = ctx->ns_steps;
  if (refactors) *refactors = ctx->ns_refactors;
  return 0;
}

int mgpbd_mg_set_coarse_update(mgpbd_ctx* ctx, int newton_schulz) {
  CTX_CHECK("mgpbd_mg_set_coarse_update");
  ctx->ns_update = newton_schulz != 0;
  return 0;
}

// V-cycle on level l >= 0 for rhs b with the first pre-smooth sweep (x = omega D^-1 b, or
// x = 0 when n_smooth == 0) already in x_first (one of the level's x buffers). Returns the
// buffer holding the result.
static float* vcycle_from(mgpbd_ctx* ctx, int l, const float* b, float* x_first, const int* done) {
  MgLevel& L = ctx->L[l];
  const bool coarsest = (l == ctx->n_levels - 1);
  if (coarsest) {
    if (ctx->coarse_mode == MGPBD_COARSE_CHOLESKY) {
      k_dense_matvec<<<(L.n + 3) / 4, kBlock, 0, ctx->stream>>>(L.n, ctx->dense, b, x_first, done);
      return x_first;
    }
    int total = std::min(ctx->coarse_sweeps, L.n);
    return sweeps(ctx, l, b, x_first, total - 1, done);
  }
  float* x = sweeps(ctx, l, b, x_first, ns(ctx, l) - 1, done, 2, true);
  float* r = L.wb(3);
  MgLevel& C = ctx->L[l + 1];
  // residual r = b - A x (x == 0 when n_smooth == 0: then r = b)
  if (ctx->n_smooth > 0) {
    int pe = prof_begin(ctx, l == 0 ? MGPBD_PROF_RESID_L0 : MGPBD_PROF_COARSE);
    sell_launch<2>(ctx, L, x, r, nullptr, b, 0.f, nullptr, done);
    prof_end(ctx, pe);
  } else {
    k_blas1<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(2, L.n, 0.f, b, r);
  }
  // restrict + the coarse level's first sweep from zero
  float* bc = C.wb(0);
  float* xc = C.wb(1);
  const bool c_coarsest = (l + 1 == ctx->n_levels - 1);
  const bool c_dense = c_coarsest && ctx->coarse_mode == MGPBD_COARSE_CHOLESKY;
  const bool c_first = !c_dense && (c_coarsest || ctx->n_smooth > 0);
  k_restrict<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(C.n, L.mptr, L.mem, r, bc,
                                                       C.dinv, c_first ? xc : nullptr,
                                                       ctx->omega, wl(ctx, l + 1), done);
  if (!c_first && !c_dense) k_blas1<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(3, C.n, 0.f, nullptr, xc);
  float* xcr = vcycle_from(ctx, l + 1, bc, xc, done);
  if (ctx->cycle == 2 && !c_coarsest) {
    // W-cycle: a second coarse cycle on the coarse residual, added to the first result
    float* save = C.wb(5);
    float* rc = C.wb(4);
    k_copy_d<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(C.n, xcr, save, done);
    sell_launch<2>(ctx, C, xcr, rc, nullptr, bc, 0.f, nullptr, done);
    float* x2 = C.wb(1);
    if (ctx->n_smooth > 0)
      k_first_sweep<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(C.n, C.dinv, rc, x2, ctx->omega,
                                                                 wl(ctx, l + 1), done);
    else
      k_blas1<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(3, C.n, 0.f, nullptr, x2);
    xcr = vcycle_from(ctx, l + 1, rc, x2, done);
    k_add_d<<<nblocks(C.n), kBlock, 0, ctx->stream>>>(C.n, save, xcr, done);
  }
  k_prolong<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.agg, xcr, x, ctx->overcorrect, done);
  return sweeps(ctx, l, b, x, ns(ctx, l), done, 1, false);
}

// ---------------- reduced-precision (scaled) V-cycle ----------------
extern "C++" {
// Calls f with the row-vertex source of the factored level 0's tet passes: the packed words
// when mgpbd_mg_set_level0_packed_rows registered them, else the int4 list.
template <class Fn>
static void with_rows(const mgpbd_ctx::Factored& F, Fn&& f) {
  if (F.tpk) f(RowsPk{F.tpk, F.pk_bb, F.pk_ob});
  else f(RowsI4{F.t_i});
}

template <int VT>
static void launch_sellh(mgpbd_ctx* ctx, const MgLevel& L, int mode, const float* x, float* y,
                         const float* b, const int* done, const float* wp, const float* xp,
                         const float* cc) {
  const int g = nblocks(64LL * L.n_slices);
  hipStream_t s = ctx->stream;
  const float w = ctx->omega;
#define SELLH(D, M) k_sellh<VT, D, M><<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, \
                                                         L.hpk, L.hval, x, y, b, w, wp, L.isd, done, xp, cc)
#define SELLW(M) k_sellw<VT, M><<<g, kBlock, lds, s>>>(L.n, L.n_slices, L.slice_ptr, L.hpk, L.seg_ptr, \
                                                     L.segs, x, y, b, w, wp, L.isd, done, xp, cc)
  const size_t lds = sizeof(float) * kWinSeg * L.win_segs;
  if (L.hpk && L.segs) {
    switch (mode) {
      case 1: SELLW(1); break;
      case 2: SELLW(2); break;
      case 3: SELLW(3); break;
      case 4: SELLW(4); break;
      default: SELLW(5); break;
    }
  } else if (L.hpk) {
    switch (mode) {
      case 1: SELLH(true, 1); break;
      case 2: SELLH(true, 2); break;
      case 3: SELLH(true, 3); break;
      case 4: SELLH(true, 4); break;
      default: SELLH(true, 5); break;
    }
  } else {
    switch (mode) {
      case 1: SELLH(false, 1); break;
      case 2: SELLH(false, 2); break;
      case 3: SELLH(false, 3); break;
      case 4: SELLH(false, 4); break;
      default: SELLH(false, 5); break;
    }
  }
#undef SELLH
#undef SELLW
}

// The factored level's vertex pass u = K^T x on the context stream: the uniform-layout kernel
// when the layout is uniform, else the per-slice one (same sums, bit for bit).
template <int VT>
static void fac_gather(mgpbd_ctx* ctx, const float* x, const int* done) {
  const mgpbd_ctx::Factored& F = ctx->fac;
  hipStream_t s = ctx->stream;
  const int g = nblocks((int64_t)kFacLanes * F.nv);
  switch (F.tbw ? F.rows_u : 0) {
#define FGU(R) case R: k_fac_gather_u<VT, R><<<g, kBlock, 0, s>>>(F.nv, F.tbw, F.vs_k, x, F.u, done); return;
    FGU(1) FGU(2) FGU(3) FGU(4) FGU(5) FGU(6) FGU(7) FGU(8)
#undef FGU
    default:
      k_fac_gather<VT><<<g, kBlock, 0, s>>>(F.nv, F.vs_ptr, F.vs_cnt, F.vs_inc, F.tbase, F.vs_k,
                                            F.invm, F.ic2, x, F.u, done);
  }
}

template <int VT>
static void launch_fac(mgpbd_ctx* ctx, int mode, const float* x, float* y, const float* b,
                       const int* done, const float* wp, const float* xp, const float* cc) {
  const mgpbd_ctx::Factored& F = ctx->fac;
  const MgLevel& L = ctx->L[0];
  hipStream_t s = ctx->stream;
  fac_gather<VT>(ctx, x, done);
  const int g = nblocks(L.n);
  const float w = ctx->omega;
  with_rows(F, [&](auto rows) {
    using RS = decltype(rows);
#define FAC(M) k_fac_apply<VT, M, RS><<<g, kBlock, 0, s>>>(L.n, rows, F.k(), F.ahat, F.u, x, y, b, \
                                                           w, wp, L.isd, done, xp, cc)
    switch (mode) {
      case 1: FAC(1); break;
      case 2: FAC(2); break;
      case 3: FAC(3); break;
      case 4: FAC(4); break;
      default: FAC(5); break;
    }
#undef FAC
  });
}
}  // extern "C++"

static void sellh(mgpbd_ctx* ctx, const MgLevel& L, int mode, const float* x, float* y,
                  const float* b, const int* done, const float* wp, const float* xp = nullptr,
                  const float* cc = nullptr) {
  if (&L == &ctx->L[0] && ctx->fac.on) {
    if (ctx->precision == MGPBD_PREC_FP16) launch_fac<1>(ctx, mode, x, y, b, done, wp, xp, cc);
    else launch_fac<2>(ctx, mode, x, y, b, done, wp, xp, cc);
    return;
  }
  if (ctx->precision == MGPBD_PREC_FP16) launch_sellh<1>(ctx, L, mode, x, y, b, done, wp, xp, cc);
  else launch_sellh<2>(ctx, L, mode, x, y, b, done, wp, xp, cc);
}

// `count` scaled sweeps from cur; with z_out the last one writes isd * x into z_out instead.
// Scaled-level counterpart of `sweeps` (same step numbering); the last step of a phase that
// ends the V-cycle on level 0 writes the unscaled result into z_out.
static float* sweeps_h(mgpbd_ctx* ctx, int l, const float* b, float* cur, int count,
                       const int* done, float* z_out, int first_step = 0, bool from_zero = false) {
  MgLevel& L = ctx->L[l];
  float* xa = L.wb(1);
  float* xb = L.wb(2);
  const bool cheb = first_step > 0 && cheb_level(ctx, l);
  for (int sw = 0; sw < count; ++sw) {
    const bool last_out = z_out && sw == count - 1;
    float* other = (cur == xa) ? xb : xa;   // holds the previous iterate
    float* nxt = last_out ? z_out : other;
    const int step = first_step + sw;
    int pe = prof_begin(ctx, l == 0 ? MGPBD_PROF_JACOBI_L0 : MGPBD_PROF_COARSE);
    if (cheb && step >= 2) {
      const float* xp = (step == 2 && from_zero) ? nullptr : other;
      sellh(ctx, L, last_out ? 5 : 4, cur, nxt, b, done, nullptr, xp, cl(ctx, l, step));
    } else {
      sellh(ctx, L, last_out ? 3 : 1, cur, nxt, b, done, wl(ctx, l));
    }
    prof_end(ctx, pe);
    cur = nxt;
  }
  return cur;
}

// Scaled V-cycle on level l for rhs b^ with the first sweep (x^ = omega b^) in x_first.
// On level 0 the result is returned unscaled in z_out. On the dense coarsest level, push.x set
// makes the product also prolong onto the level above (the caller then skips k_prolong_h).
static float* vcycle_h(mgpbd_ctx* ctx, int l, const float* b, float* x_first, const int* done,
                       float* z_out, Push push = Push{}) {
  MgLevel& L = ctx->L[l];
  hipStream_t s = ctx->stream;
  const bool coarsest = (l == ctx->n_levels - 1);
  if (coarsest) {
    if (ctx->coarse_mode == MGPBD_COARSE_CHOLESKY) {
      k_dense_matvec_scaled<<<(L.n + 3) / 4, kBlock, 0, s>>>(L.n, ctx->dense, L.sd, b, x_first,
                                                             done, push);
      return x_first;
    }
    const int total = std::min(ctx->coarse_sweeps, L.n);
    return sweeps_h(ctx, l, b, x_first, total - 1, done, nullptr);
  }
  float* x = sweeps_h(ctx, l, b, x_first, ns(ctx, l) - 1, done, nullptr, 2, true);
  float* r = L.wb(3);
  int pe = prof_begin(ctx, l == 0 ? MGPBD_PROF_RESID_L0 : MGPBD_PROF_COARSE);
  sellh(ctx, L, 2, x, r, b, done, nullptr);
  prof_end(ctx, pe);
  MgLevel& C = ctx->L[l + 1];
  float* bc = C.wb(0);
  float* xc = C.wb(1);
  const bool c_dense = (l + 1 == ctx->n_levels - 1) && ctx->coarse_mode == MGPBD_COARSE_CHOLESKY;
  k_restrict_h<<<nblocks(C.n), kBlock, 0, s>>>(C.n, L.mptr, L.mem, L.sd, r, C.isd, bc,
                                               c_dense ? nullptr : xc, ctx->omega, wl(ctx, l + 1),
                                               done);
  // the dense coarsest product prolongs onto x itself (no W-cycle doubling happens there)
  const bool w2 = ctx->cycle == 2 && l + 1 < ctx->n_levels - 1;
  const bool pushed = c_dense && ctx->fused_launch;
  const Push pc = pushed ? Push{L.mptr, L.mem, L.sd, C.isd, x, ctx->overcorrect} : Push{};
  float* xcr = vcycle_h(ctx, l + 1, bc, xc, done, nullptr, pc);
  if (w2) {
    float* save = C.wb(5);
    float* rc = C.wb(4);
    k_copy_d<<<nblocks(C.n), kBlock, 0, s>>>(C.n, xcr, save, done);
    sellh(ctx, C, 2, xcr, rc, bc, done, nullptr);
    float* x2 = C.wb(1);
    k_scale_d<<<nblocks(C.n), kBlock, 0, s>>>(C.n, ctx->omega, wl(ctx, l + 1), rc, x2, done);
    xcr = vcycle_h(ctx, l + 1, rc, x2, done, nullptr);
    k_add_d<<<nblocks(C.n), kBlock, 0, s>>>(C.n, save, xcr, done);
  }
  if (!pushed)
    k_prolong_h<<<nblocks(L.n), kBlock, 0, s>>>(L.n, L.agg, L.sd, C.isd, xcr, x,
                                                ctx->overcorrect, done);
  return sweeps_h(ctx, l, b, x, ns(ctx, l), done, z_out, 1, false);
}

int mgpbd_mg_set_level_smoothing(mgpbd_ctx* ctx, int level, int steps) {
  CTX_CHECK("mgpbd_mg_set_level_smoothing");
  // steps 0 restores the configured n_smooth; per-level counts need n_smooth >= 1 (the
  // first step of every phase structure is laid out for a smoothed cycle)
  if (level < 0 || level >= ctx->n_levels || steps < 0 || steps > kChebSteps ||
      (steps > 0 && ctx->n_smooth < 1))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_smoothing: level in range, steps in "
                                    "[0, 16], n_smooth >= 1");
  ctx->level_smooth[level] = steps;
  drop_graphs(ctx);
  return 0;
}

int mgpbd_mg_set_cycle(mgpbd_ctx* ctx, int cycle, float overcorrect) {
  CTX_CHECK("mgpbd_mg_set_cycle");
  if ((cycle != 1 && cycle != 2) || !(overcorrect > 0.0f && overcorrect < 4.0f))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_cycle: cycle 1 (V) or 2 (W), overcorrect in (0, 4)");
  ctx->cycle = cycle;
  ctx->overcorrect = overcorrect;
  drop_graphs(ctx);
  return 0;
}

int mgpbd_mg_set_launch_fusion(mgpbd_ctx* ctx, int enable) {
  CTX_CHECK("mgpbd_mg_set_launch_fusion");
  ctx->fused_launch = enable != 0;
  drop_graphs(ctx);   // captured replays hold the launches of the previous choice
  return 0;
}

int mgpbd_mg_set_smoother(mgpbd_ctx* ctx, int mode, float target, float safety, int power_iters,
                          float cheb_ratio) {
  CTX_CHECK("mgpbd_mg_set_smoother");
  if (mode < 0 || mode > 2 || !(target > 0.0f && target < 2.0f) || !(safety >= 1.0f) ||
      power_iters < 1 || power_iters > 64 || !(cheb_ratio > 1.0f && cheb_ratio < 1e4f))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_smoother: mode 0/1/2, target in (0, 2), "
                                    "safety >= 1, power_iters in [1, 64], cheb_ratio > 1");
  ctx->spectral = mode >= 1;
  ctx->fixed_weights = false;
  ctx->chebyshev = mode == 2;
  ctx->cheb_ratio = cheb_ratio;
  ctx->sm_target = target;
  ctx->sm_safety = safety;
  ctx->power_iters = power_iters;
  for (int l = 0; l < MGPBD_MAX_LEVELS; ++l) {
    ctx->pv_ready[l] = false;
    k_set_float<<<1, 64, 0, ctx->stream>>>(ctx->omega_dev + l, ctx->omega);
  }
  drop_graphs(ctx);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_set_smoother");
  return 0;
}

// Refresh of the spectral estimates and the weights/coefficients derived from them. fresh:
// restart every level's power iteration from the deterministic start vector with 12 steps
// (the stall rescue of the PCG); boost multiplies the safety factor of this refresh.
static int smoother_refresh(mgpbd_ctx* ctx, bool fresh, float boost) {
  hipStream_t s = ctx->stream;
  for (int l = 0; l < ctx->n_levels; ++l) {
    MgLevel& L = ctx->L[l];
    const bool coarsest = l == ctx->n_levels - 1;
    if (coarsest && ctx->coarse_mode == MGPBD_COARSE_CHOLESKY) continue;
    if (L.diagonal) {   // D^-1 A = I: Jacobi with weight 1 is exact
      k_set_float<<<1, 64, 0, s>>>(ctx->omega_dev + l, 1.0f);
      continue;
    }
    if (nblocks(64LL * L.n_slices) > kPartialStride)
      return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_update_smoother: level exceeds the reduction capacity");
    float* v = L.wb(6);
    float* w = L.wb(7);
    int iters = ctx->power_iters;
    if (!ctx->pv_ready[l] || fresh) {
      k_power_init<<<nblocks(L.n), kBlock, 0, s>>>(L.n, v);
      if (!ctx->pv_ready[l]) ctx->pv_updates[l] = 0;
      ctx->pv_ready[l] = true;
    }
    // the first updates after a (re)set start far from the top eigenvector (at rest the
    // level-0 matrix is diagonal): longer runs until the vector has followed the matrix; also
    // after an update whose last two estimates still differed by more than 2 % (read from the
    // pinned copy of that update, which the Newton loop's own synchronisations have completed)
    const double* ph = ctx->pw_host + 4 * l;
    const bool moving = !(std::fabs(ph[0] - ph[2]) <= 0.02 * ph[0]);
    if (ctx->pv_updates[l] < 4 || fresh) iters = std::max(iters, 12);
    else if (moving) iters = std::max(iters, 8);
    // else power_iters steps: the settled vector follows the matrix from refresh to refresh
    ++ctx->pv_updates[l];
    const int g = nblocks(64LL * L.n_slices);
    double* pw = ctx->pw_dev + 4 * l;
    const mgpbd_ctx::Factored& F = ctx->fac;
    const bool fac = l == 0 && F.on && ctx->precision != MGPBD_PREC_FP32;
    for (int it = 0; it < iters; ++it) {
      if (fac && ctx->precision == MGPBD_PREC_FP16) {
        fac_gather<1>(ctx, v, nullptr);
        with_rows(F, [&](auto rows) {
          k_fac_power<1, decltype(rows)><<<g, kBlock, 0, s>>>(L.n, rows, F.k(), F.ahat, F.u, v, w,
                                                              ctx->ws);
        });
      } else if (fac) {
        fac_gather<2>(ctx, v, nullptr);
        with_rows(F, [&](auto rows) {
          k_fac_power<2, decltype(rows)><<<g, kBlock, 0, s>>>(L.n, rows, F.k(), F.ahat, F.u, v, w,
                                                              ctx->ws);
        });
      } else if (ctx->precision == MGPBD_PREC_FP32 && L.woff)
        k_power_step_wf<<<g, kBlock, sizeof(float) * kWinSeg * L.win_segs, s>>>(
            L.n, L.n_slices, L.slice_ptr, L.woff, L.val, L.seg_ptr, L.segs, L.dinv, v, w, ctx->ws);
      else if (ctx->precision == MGPBD_PREC_FP32)
        k_power_step<<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.val, L.dinv, v,
                                          w, ctx->ws);
      else if (L.segs && ctx->precision == MGPBD_PREC_FP16)
        k_power_step_w<1><<<g, kBlock, sizeof(float) * kWinSeg * L.win_segs, s>>>(
            L.n, L.n_slices, L.slice_ptr, L.hpk, L.seg_ptr, L.segs, v, w, ctx->ws);
      else if (L.segs)
        k_power_step_w<2><<<g, kBlock, sizeof(float) * kWinSeg * L.win_segs, s>>>(
            L.n, L.n_slices, L.slice_ptr, L.hpk, L.seg_ptr, L.segs, v, w, ctx->ws);
      else if (ctx->precision == MGPBD_PREC_FP16 && L.hpk)
        k_power_step_h<1, true><<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.hpk,
                                                     L.hval, v, w, ctx->ws);
      else if (ctx->precision == MGPBD_PREC_FP16)
        k_power_step_h<1, false><<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.hpk,
                                                      L.hval, v, w, ctx->ws);
      else if (L.hpk)
        k_power_step_h<2, true><<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.hpk,
                                                     L.hval, v, w, ctx->ws);
      else
        k_power_step_h<2, false><<<g, kBlock, 0, s>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.hpk,
                                                      L.hval, v, w, ctx->ws);
      if (g <= 1024) {   // every block re-sums the (few) partials: one launch instead of two
        k_power_fin_norm<<<nblocks(L.n), kBlock, 0, s>>>(L.n, ctx->ws, g, w, v, pw);
      } else {
        k_power_fin<<<1, kBlock, 0, s>>>(ctx->ws, g, pw);
        k_power_norm<<<nblocks(L.n), kBlock, 0, s>>>(L.n, w, v, pw);
      }
    }
    k_power_omega<<<1, 64, 0, s>>>(pw, ctx->sm_target, ctx->sm_safety * boost, ctx->omega,
                                   cheb_level(ctx, l) ? 1 : 0, ctx->cheb_ratio,
                                   ctx->omega_dev + l, ctx->cheb_dev + 2 * kChebSteps * l);
  }
  MGPBD_LAUNCH_CHECK("mgpbd_mg_update_smoother");
  MGPBD_HIP_TRY(hipMemcpyAsync(ctx->pw_host, ctx->pw_dev, sizeof(double) * 4 * MGPBD_MAX_LEVELS,
                               hipMemcpyDeviceToHost, s));
  return 0;
}

int mgpbd_mg_set_level_weights(mgpbd_ctx* ctx, int n, const float* omega) {
  CTX_CHECK("mgpbd_mg_set_level_weights");
  if (n < 1 || !omega) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_weights: n >= 1 weights");
  for (int i = 0; i < n; ++i)
    if (!(omega[i] > 0.0f && omega[i] < 2.0f))
      return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_weights: weights in (0, 2)");
  ctx->spectral = true;        // the sweeps read their weight from omega_dev
  ctx->fixed_weights = true;
  ctx->chebyshev = false;
  for (int l = 0; l < MGPBD_MAX_LEVELS; ++l)
    k_set_float<<<1, 64, 0, ctx->stream>>>(ctx->omega_dev + l, omega[std::min(l, n - 1)]);
  drop_graphs(ctx);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_set_level_weights");
  return 0;
}

int mgpbd_mg_update_smoother(mgpbd_ctx* ctx, double* lambda_out) {
  CTX_CHECK("mgpbd_mg_update_smoother");
  if (int e = check_levels(ctx)) return e;
  if (!ctx->spectral || ctx->fixed_weights) return 0;
  hipStream_t s = ctx->stream;
  if (int e = smoother_refresh(ctx, false, 1.0f)) return e;
  if (lambda_out) {
    double tmp[4 * MGPBD_MAX_LEVELS];
    MGPBD_HIP_TRY(hipMemcpyAsync(tmp, ctx->pw_dev, sizeof(tmp), hipMemcpyDeviceToHost, s));
    MGPBD_HIP_TRY(hipStreamSynchronize(s));
    for (int l = 0; l < ctx->n_levels; ++l) lambda_out[l] = tmp[4 * l];
  }
  return 0;
}

int mgpbd_mg_get_smoother(mgpbd_ctx* ctx, float* omega_out, float* cheb_out) {
  CTX_CHECK("mgpbd_mg_get_smoother");
  if (int e = check_levels(ctx)) return e;
  if (!omega_out) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_get_smoother: null output");
  if (!ctx->spectral) {
    for (int l = 0; l < ctx->n_levels; ++l) omega_out[l] = ctx->omega;
    if (cheb_out) std::memset(cheb_out, 0, sizeof(float) * 2 * kChebSteps * ctx->n_levels);
    return 0;
  }
  float tmp[MGPBD_MAX_LEVELS];
  MGPBD_HIP_TRY(hipMemcpyAsync(tmp, ctx->omega_dev, sizeof(tmp), hipMemcpyDeviceToHost, ctx->stream));
  if (cheb_out)
    MGPBD_HIP_TRY(hipMemcpyAsync(cheb_out, ctx->cheb_dev, sizeof(float) * 2 * kChebSteps * ctx->n_levels,
                                 hipMemcpyDeviceToHost, ctx->stream));
  MGPBD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int l = 0; l < ctx->n_levels; ++l) omega_out[l] = tmp[l];
  if (cheb_out && !ctx->chebyshev)
    std::memset(cheb_out, 0, sizeof(float) * 2 * kChebSteps * ctx->n_levels);
  return 0;
}

int mgpbd_mg_set_precision(mgpbd_ctx* ctx, int precision, int flexible) {
  CTX_CHECK("mgpbd_mg_set_precision");
  if (precision != MGPBD_PREC_FP32 && precision != MGPBD_PREC_FP16 && precision != MGPBD_PREC_BF16)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_precision: unknown precision");
  if (precision != MGPBD_PREC_FP32 && ctx->n_smooth < 1)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_precision: the reduced-precision V-cycle needs n_smooth >= 1");
  ctx->precision = precision;
  ctx->flexible = flexible != 0;
  drop_graphs(ctx);
  return 0;
}

int mgpbd_mg_set_level_half(mgpbd_ctx* ctx, int level, uint16_t* hval, uint32_t* hpacked,
                            float* sd, float* isd, const int32_t* diag_pos) {
  CTX_CHECK("mgpbd_mg_set_level_half");
  if (level < 0 || level >= ctx->n_levels || (hval && hpacked) || (!hval && !hpacked && level != 0) ||
      !sd || !isd || !diag_pos)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_half: invalid arguments");
  MgLevel& L = ctx->L[level];
  L.hval = hpacked ? nullptr : hval; L.hpk = hpacked; L.sd = sd; L.isd = isd; L.diag_pos = diag_pos;
  L.seg_ptr = nullptr; L.segs = nullptr; L.win_segs = 0; L.woff = nullptr;
  drop_graphs(ctx);
  return 0;
}

int mgpbd_mg_set_level_window(mgpbd_ctx* ctx, int level, const int32_t* seg_ptr,
                              const int32_t* segs, int n_blocks, int max_segs, uint16_t* woff) {
  CTX_CHECK("mgpbd_mg_set_level_window");
  if (level < 0 || level >= ctx->n_levels || !seg_ptr || !segs)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_window: invalid arguments");
  MgLevel& L = ctx->L[level];
  if (!L.idx || L.n <= 0 || (!L.hpk && !woff))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_window: level needs packed half words "
                                    "(mgpbd_mg_set_level_half with hpacked) or woff");
  if (n_blocks != nblocks(64LL * L.n_slices) || max_segs < 1 || max_segs > kWinMaxSegs)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_window: n_blocks must be ceil(64 n_slices / " +
                                    std::to_string(kBlock) + "), max_segs in [1, " +
                                    std::to_string(kWinMaxSegs) + "]");
  int* bad = reinterpret_cast<int*>(ctx->ws.partials + (size_t)3 * kPartialStride);
  MGPBD_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  k_window_offsets<<<n_blocks, kBlock, 0, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.idx,
                                                           seg_ptr, segs, L.hpk, woff, bad);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_set_level_window");
  int hb = 0;
  MGPBD_HIP_TRY(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MGPBD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (hb) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level_window: a column lies outside its "
                                          "block's window");
  L.seg_ptr = seg_ptr; L.segs = segs; L.win_segs = max_segs; L.woff = woff;
  drop_graphs(ctx);
  return 0;
}

int mgpbd_mg_set_level0_factored(mgpbd_ctx* ctx, int n_verts, const int32_t* t_i,
                                 const float* v_invm, const float* gradC,
                                 const float* alpha_tilde, const int32_t* vs_ptr,
                                 const int32_t* vs_cnt, const int32_t* vs_inc, uint32_t* vs_k,
                                 float* vs_g, uint32_t* tk, float* ahat, float* u, double* u64,
                                 float* dinv) {
  CTX_CHECK("mgpbd_mg_set_level0_factored");
  if (ctx->n_levels < 1 || ctx->L[0].n <= 0)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_factored: level 0 not set");
  drop_graphs(ctx);
  if (!t_i) {   // back to the level's stored half values
    ctx->fac = mgpbd_ctx::Factored{};
    return 0;
  }
  if (n_verts <= 0 || !v_invm || !gradC || !alpha_tilde || !vs_ptr || !vs_cnt || !vs_inc ||
      !vs_k || !vs_g || !tk || !ahat || !u || !u64 || !dinv)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_factored: invalid arguments");
  const auto al = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; };
  if (!al(t_i, 16) || !al(gradC, 16) || !al(vs_k, 8) || !al(tk, 16) || !al(u, 16) ||
      !al(vs_g, 16) || !al(u64, 32))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_factored: t_i, gradC, vs_g, tk and u "
                                    "need 16-byte, u64 32-byte, vs_k 8-byte alignment");
  if (!ctx->L[0].sd || !ctx->L[0].isd)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_factored: call mgpbd_mg_set_level_half "
                                    "for level 0 first (its scales)");
  // K's scale c: c^2 = the largest inverse mass (read once at set-up, after whatever stream
  // wrote the masses; they are the mesh's)
  std::vector<float> w(n_verts);
  MGPBD_HIP_TRY(hipDeviceSynchronize());
  MGPBD_HIP_TRY(hipMemcpy(w.data(), v_invm, sizeof(float) * (size_t)n_verts,
                          hipMemcpyDeviceToHost));
  float wmax = 0.0f;
  for (float x : w)
    if (std::isfinite(x)) wmax = std::max(wmax, x);
  mgpbd_ctx::Factored& F = ctx->fac;
  F.cs = wmax > 0.0f ? std::sqrt(wmax) : 1.0f;
  F.ic2 = 1.0f / (F.cs * F.cs);
  F.nv = n_verts;
  F.t_i = reinterpret_cast<const int4*>(t_i);
  F.invm = v_invm; F.gradC = gradC; F.alpha = alpha_tilde;
  F.vs_ptr = vs_ptr; F.vs_cnt = vs_cnt; F.vs_inc = vs_inc;
  F.vs_k = reinterpret_cast<v2u32*>(vs_k);
  F.vs_g = reinterpret_cast<float4*>(vs_g);
  const int nt = ctx->L[0].n;
  F.tka = reinterpret_cast<v4u32*>(tk);
  F.tkz = reinterpret_cast<unsigned short*>(tk + 4 * (size_t)nt);
  F.ahat = ahat;
  F.u = reinterpret_cast<fvec3*>(u);
  F.u64 = reinterpret_cast<double4*>(u64);
  F.dinv = dinv;
  F.galerkin = false;
  F.on = true;
  return 0;
}

int mgpbd_mg_set_level0_offsets(mgpbd_ctx* ctx, const int32_t* tbase, int rows_uniform) {
  CTX_CHECK("mgpbd_mg_set_level0_offsets");
  mgpbd_ctx::Factored& F = ctx->fac;
  if (!F.on) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_offsets: no factored level 0");
  if (rows_uniform < 0 || rows_uniform > 8 || (rows_uniform > 0 && !tbase))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_offsets: rows_uniform in [0, 8], and "
                                    "a uniform layout needs the base tets");
  drop_graphs(ctx);
  if (tbase && ctx->fac_tbw_cap < F.nv) {
    if (ctx->fac_tbw) MGPBD_HIP_TRY(hipFree(ctx->fac_tbw));
    ctx->fac_tbw = nullptr;
    ctx->fac_tbw_cap = 0;
    MGPBD_HIP_TRY(hipMalloc(&ctx->fac_tbw, sizeof(int2) * (size_t)F.nv));
    ctx->fac_tbw_cap = F.nv;
  }
  F.tbase = tbase;
  F.tbw = tbase ? ctx->fac_tbw : nullptr;
  F.rows_u = rows_uniform;
  return 0;
}

int mgpbd_mg_set_level0_packed_rows(mgpbd_ctx* ctx, const uint64_t* rows_dev, int base_bits,
                                    int diff_bits) {
  CTX_CHECK("mgpbd_mg_set_level0_packed_rows");
  mgpbd_ctx::Factored& F = ctx->fac;
  if (!F.on) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_packed_rows: no factored level 0");
  if (rows_dev && (base_bits < 1 || base_bits > 31 || diff_bits < 2 || diff_bits > 20 ||
                   base_bits + 3 * diff_bits > 64))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level0_packed_rows: base_bits in [1, 31], "
                                    "diff_bits in [2, 20], base_bits + 3 diff_bits <= 64");
  drop_graphs(ctx);
  F.tpk = rows_dev;
  F.pk_bb = rows_dev ? base_bits : 0;
  F.pk_ob = rows_dev ? diff_bits : 0;
  return 0;
}

int mgpbd_mg_set_level1_factored(mgpbd_ctx* ctx, int n_pairs, const int32_t* gptr,
                                 const int32_t* gmem, const int32_t* gvert, float* G,
                                 int n_entries, const int32_t* tptr, const int32_t* terms,
                                 const int32_t* cpos, const int32_t* cpos_t,
                                 const int32_t* cagg, float* c_val) {
  CTX_CHECK("mgpbd_mg_set_level1_factored");
  mgpbd_ctx::Factored& F = ctx->fac;
  if (!F.on || ctx->n_levels < 2 || !ctx->L[0].mptr)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level1_factored: needs the factored level 0 "
                                    "(mgpbd_mg_set_level0_factored) and a level 1");
  if (!gptr) {
    F.galerkin = false;
    return 0;
  }
  if (n_pairs <= 0 || n_entries <= 0 || !gmem || !gvert || !G || !tptr || !terms || !cpos ||
      !cpos_t || !cagg || !c_val || (reinterpret_cast<uintptr_t>(G) % 16) != 0 ||
      (reinterpret_cast<uintptr_t>(terms) % 8) != 0)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_set_level1_factored: invalid arguments (G 16-byte, "
                                    "terms 8-byte aligned)");
  F.ng = n_pairs; F.gptr = gptr; F.gmem = gmem; F.gvert = gvert;
  F.G = reinterpret_cast<float4*>(G);
  F.ne = n_entries; F.tptr = tptr; F.terms = reinterpret_cast<const int2*>(terms);
  F.cpos = cpos; F.cpos_t = cpos_t; F.cagg = cagg; F.cval = c_val;
  F.galerkin = true;
  return 0;
}

int mgpbd_mg_factored_scatter(mgpbd_ctx* ctx, const float* dlam, float* lambda, float* dpos) {
  CTX_CHECK("mgpbd_mg_factored_scatter");
  const mgpbd_ctx::Factored& F = ctx->fac;
  if (!F.on || !dlam || !lambda || !dpos)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_factored_scatter: needs the factored level 0 and "
                                    "device arrays");
  hipStream_t s = ctx->stream;
  const int n = ctx->L[0].n;
  k_add_inplace<<<nblocks(n), kBlock, 0, s>>>(n, dlam, lambda);
  k_fac_scatter<<<nblocks((int64_t)kFacLanes * F.nv), kBlock, 0, s>>>(
      F.nv, F.vs_ptr, F.vs_cnt, F.vs_inc, F.vs_g, F.invm, dlam, dpos);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_factored_scatter");
  return 0;
}

int mgpbd_mg_factored_refill(mgpbd_ctx* ctx) {
  CTX_CHECK("mgpbd_mg_factored_refill");
  const mgpbd_ctx::Factored& F = ctx->fac;
  if (!F.on || !F.galerkin)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_factored_refill: call mgpbd_mg_set_level1_factored first");
  hipStream_t s = ctx->stream;
  const MgLevel& L = ctx->L[0];
  const int pe = prof_begin(ctx, MGPBD_PROF_REFILL);
  k_fac_gsum<<<nblocks(F.ng), kBlock, 0, s>>>(F.ng, F.gptr, F.gmem, F.gvert, F.gradC, F.invm, F.G);
  k_fac_galerkin<<<nblocks(F.ne), kBlock, 0, s>>>(F.ne, F.tptr, F.terms, F.G, F.cpos, F.cpos_t,
                                                  F.cagg, L.mptr, L.mem, F.alpha, F.cval);
  prof_end(ctx, pe);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_factored_refill");
  return 0;
}

static int convert_half_levels(mgpbd_ctx* ctx, int first, int last);

int mgpbd_mg_convert_half(mgpbd_ctx* ctx) {
  CTX_CHECK("mgpbd_mg_convert_half");
  if (int e = check_levels(ctx)) return e;
  return convert_half_levels(ctx, 0, ctx->n_levels);
}

int mgpbd_mg_convert_half_levels(mgpbd_ctx* ctx, int first, int last) {
  CTX_CHECK("mgpbd_mg_convert_half_levels");
  if (int e = check_levels(ctx)) return e;
  if (first < 0 || last > ctx->n_levels || first > last)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_convert_half_levels: 0 <= first <= last <= n_levels");
  return convert_half_levels(ctx, first, last);
}

static int convert_half_levels(mgpbd_ctx* ctx, int first, int last) {
  if (ctx->precision == MGPBD_PREC_FP32) return 0;
  for (int l = first; l < last; ++l) {
    MgLevel& L = ctx->L[l];
    const bool fac = l == 0 && ctx->fac.on;
    if (!L.hval && !L.hpk && !fac) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_convert_half: level " + std::to_string(l) + " has no half buffers");
    if (fac) {   // scales and K from the gradients (level 0 is not assembled per iteration)
      const mgpbd_ctx::Factored& F = ctx->fac;
      const int gv = nblocks((int64_t)kFacLanes * F.nv);
      hipStream_t s = ctx->stream;
      const int gt = nblocks(L.n);
      if (ctx->precision == MGPBD_PREC_FP16) {
        k_fac_tet<1><<<gt, kBlock, 0, s>>>(L.n, F.t_i, F.invm, F.gradC, F.alpha, F.cs, L.sd,
                                           L.isd, F.dinv, F.tka, F.tkz, F.ahat);
        k_fac_vert<1><<<gv, kBlock, 0, s>>>(F.nv, F.vs_ptr, F.vs_cnt, F.vs_inc, F.tbase,
                                            F.gradC, F.k(), F.vs_k, F.vs_g, F.invm, F.ic2,
                                            F.tbw);
      } else {
        k_fac_tet<2><<<gt, kBlock, 0, s>>>(L.n, F.t_i, F.invm, F.gradC, F.alpha, F.cs, L.sd,
                                           L.isd, F.dinv, F.tka, F.tkz, F.ahat);
        k_fac_vert<2><<<gv, kBlock, 0, s>>>(F.nv, F.vs_ptr, F.vs_cnt, F.vs_inc, F.tbase,
                                            F.gradC, F.k(), F.vs_k, F.vs_g, F.invm, F.ic2,
                                            F.tbw);
      }
      continue;
    }
    k_half_scales<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.diag_pos, L.val, L.sd, L.isd);
    const int g = nblocks(64LL * L.n_slices);
    const size_t lds = sizeof(float) * kWinSeg * L.win_segs;
    if (L.woff && L.hpk && L.segs && ctx->precision == MGPBD_PREC_FP16)
      k_half_values_w<1><<<g, kBlock, lds, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.woff,
                                                         L.val, L.seg_ptr, L.segs, L.isd, L.hpk);
    else if (L.woff && L.hpk && L.segs)
      k_half_values_w<2><<<g, kBlock, lds, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.woff,
                                                         L.val, L.seg_ptr, L.segs, L.isd, L.hpk);
    else if (ctx->precision == MGPBD_PREC_FP16)
      k_half_values<1><<<g, kBlock, 0, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.val, L.isd, L.hval, L.hpk,
                                                        L.segs != nullptr);
    else
      k_half_values<2><<<g, kBlock, 0, ctx->stream>>>(L.n, L.n_slices, L.slice_ptr, L.idx, L.val, L.isd, L.hval, L.hpk,
                                                        L.segs != nullptr);
  }
  MGPBD_LAUNCH_CHECK("mgpbd_mg_convert_half");
  return 0;
}

static bool half_mode(const mgpbd_ctx* ctx) { return ctx->precision != MGPBD_PREC_FP32; }

int mgpbd_mg_vcycle(mgpbd_ctx* ctx, const float* b, float* x_out) {
  CTX_CHECK("mgpbd_mg_vcycle");
  if (int e = check_ready(ctx)) return e;
  MgLevel& L = ctx->L[0];
  if (half_mode(ctx)) {
    k_half_first<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.isd, b, L.wb(0), L.wb(1), ctx->omega,
                                                           wl(ctx, 0), nullptr);
    float* z = vcycle_h(ctx, 0, L.wb(0), L.wb(1), nullptr, L.wb(3));
    k_blas1<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(2, L.n, 0.f, z, x_out);
    MGPBD_LAUNCH_CHECK("mgpbd_mg_vcycle");
    return 0;
  }
  float* x1 = L.wb(1);
  const bool first = (ctx->n_levels == 1) || ctx->n_smooth > 0;
  if (first) k_first_sweep<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.dinv, b, x1, ctx->omega, wl(ctx, 0), nullptr);
  else k_blas1<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(3, L.n, 0.f, nullptr, x1);
  float* res = vcycle_from(ctx, 0, b, x1, nullptr);
  k_blas1<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(2, L.n, 0.f, res, x_out);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_vcycle");
  return 0;
}

int mgpbd_mg_probe_level0(mgpbd_ctx* ctx, int reps, const float* b, const float* x, float* y,
                          float* flush, int64_t flush_floats, double* ms_out) {
  CTX_CHECK("mgpbd_mg_probe_level0");
  if (int e = check_ready(ctx)) return e;
  if (reps < 1 || reps > 1000 || !b || !x || !y || x == y || !ms_out ||
      (flush_floats > 0 && !flush) || flush_floats < 0)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_probe_level0: invalid arguments");
  const MgLevel& L = ctx->L[0];
  hipStream_t s = ctx->stream;
  const bool fac = ctx->fac.on && half_mode(ctx);
  const mgpbd_ctx::Factored& F = ctx->fac;
  auto first = [&]() {   // the factored vertex pass, or a whole stored-level sweep
    if (fac) {
      if (ctx->precision == MGPBD_PREC_FP16) fac_gather<1>(ctx, x, nullptr);
      else fac_gather<2>(ctx, x, nullptr);
    } else if (half_mode(ctx)) {
      sellh(ctx, L, 1, x, y, b, nullptr, nullptr);
    } else {
      sell_launch<1>(ctx, L, x, y, L.dinv, b, ctx->omega, nullptr, nullptr);
    }
  };
  auto second = [&]() {   // the factored tet pass (nothing for a stored level)
    if (!fac) return;
    with_rows(F, [&](auto rows) {
      using RS = decltype(rows);
      if (ctx->precision == MGPBD_PREC_FP16)
        k_fac_apply<1, 1, RS><<<nblocks(L.n), kBlock, 0, s>>>(L.n, rows, F.k(), F.ahat, F.u, x, y,
                                                              b, ctx->omega, nullptr, L.isd,
                                                              nullptr, nullptr, nullptr);
      else
        k_fac_apply<2, 1, RS><<<nblocks(L.n), kBlock, 0, s>>>(L.n, rows, F.k(), F.ahat, F.u, x, y,
                                                              b, ctx->omega, nullptr, L.isd,
                                                              nullptr, nullptr, nullptr);
    });
  };
  const size_t ne = flush_floats > 0 ? 3 * (size_t)reps : 4;
  std::vector<hipEvent_t> e(ne, nullptr);
  struct Free {
    std::vector<hipEvent_t>& e;
    ~Free() { for (hipEvent_t v : e) if (v) (void)hipEventDestroy(v); }
  } fr{e};
  for (auto& ev : e) MGPBD_HIP_TRY(hipEventCreate(&ev));
  if (flush_floats == 0) {
    // warm: `reps` launches of each kind back to back between two events (per-launch time
    // including the dispatch gap, no event between kernels), then `reps` pairs
    MGPBD_HIP_TRY(hipEventRecord(e[0], s));
    for (int r = 0; r < reps; ++r) first();
    MGPBD_HIP_TRY(hipEventRecord(e[1], s));
    for (int r = 0; r < reps; ++r) second();
    MGPBD_HIP_TRY(hipEventRecord(e[2], s));
    for (int r = 0; r < reps; ++r) { first(); second(); }
    MGPBD_HIP_TRY(hipEventRecord(e[3], s));
    MGPBD_LAUNCH_CHECK("mgpbd_mg_probe_level0");
    MGPBD_HIP_TRY(hipEventSynchronize(e[3]));
    for (int k = 0; k < 3; ++k) {
      float t = 0.f;
      MGPBD_HIP_TRY(hipEventElapsedTime(&t, e[k], e[k + 1]));
      ms_out[k] = (double)t / reps;
    }
    return 0;
  }
  // flushed: a pass over a buffer larger than the caches before each rep, events around each
  // kernel (all queued before one wait)
  for (int r = 0; r < reps; ++r) {
    for (int64_t off = 0; off < flush_floats; off += (int64_t)1 << 30) {
      const int m = (int)std::min<int64_t>((int64_t)1 << 30, flush_floats - off);
      k_blas1<<<nblocks(m), kBlock, 0, s>>>(4, m, 1.0f, nullptr, flush + off);
    }
    MGPBD_HIP_TRY(hipEventRecord(e[3 * r], s));
    first();
    MGPBD_HIP_TRY(hipEventRecord(e[3 * r + 1], s));
    second();
    MGPBD_HIP_TRY(hipEventRecord(e[3 * r + 2], s));
    MGPBD_LAUNCH_CHECK("mgpbd_mg_probe_level0");
  }
  MGPBD_HIP_TRY(hipEventSynchronize(e[3 * (size_t)reps - 1]));
  double t0 = 0.0, t1 = 0.0, tp = 0.0;
  for (int r = 0; r < reps; ++r) {
    float a = 0.f, c = 0.f, d = 0.f;
    MGPBD_HIP_TRY(hipEventElapsedTime(&a, e[3 * r], e[3 * r + 1]));
    MGPBD_HIP_TRY(hipEventElapsedTime(&c, e[3 * r + 1], e[3 * r + 2]));
    MGPBD_HIP_TRY(hipEventElapsedTime(&d, e[3 * r], e[3 * r + 2]));
    t0 += a; t1 += c; tp += d;
  }
  ms_out[0] = t0 / reps;
  ms_out[1] = t1 / reps;
  ms_out[2] = tp / reps;
  return 0;
}

int mgpbd_mg_level_residual(mgpbd_ctx* ctx, int level, const float* b, const float* x, float* r) {
  CTX_CHECK("mgpbd_mg_level_residual");
  if (int e = check_ready(ctx)) return e;
  if (level < 0 || level >= ctx->n_levels || !b || !x || !r || x == r)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_mg_level_residual: invalid arguments");
  const MgLevel& L = ctx->L[level];
  if (half_mode(ctx)) sellh(ctx, L, 2, x, r, b, nullptr, nullptr);
  else sell_launch<2>(ctx, L, x, r, L.dinv, b, ctx->omega, nullptr, nullptr);
  MGPBD_LAUNCH_CHECK("mgpbd_mg_level_residual");
  return 0;
}

// The PCG operator's vertex pass (uniform-layout kernel when the layout is uniform).
extern "C++" template <typename XT>
static void fac_ogather(mgpbd_ctx* ctx, const XT* x, const int* done, const PcgState* gate) {
  const mgpbd_ctx::Factored& F = ctx->fac;
  hipStream_t s = ctx->stream;
  const int gv = nblocks((int64_t)kFacLanes * F.nv);
  switch (F.rows_u) {
#define FOU(R) case R: k_fac_ogather_u<XT, R><<<gv, kBlock, 0, s>>>(F.nv, F.tbase, F.vs_g, F.invm, x, F.u64, done, gate); return;
    FOU(1) FOU(2) FOU(3) FOU(4) FOU(5) FOU(6) FOU(7) FOU(8)
#undef FOU
    default:
      k_fac_ogather<XT><<<gv, kBlock, 0, s>>>(F.nv, F.vs_ptr, F.vs_cnt, F.vs_inc, F.tbase, F.vs_g,
                                              F.invm, x, F.u64, done, gate);
  }
}

// The PCG's level-0 operator in factored form (vertex pass, tet pass), modes of k_fac_oapply;
// x in FP32 (xf) or FP64 (xd); grid gs >= the row blocks (its partials when mode != 1).
static void fac_outer(mgpbd_ctx* ctx, int mode, const float* xf, const double* xd,
                      const float* b, float* y, const int* done, int gs,
                      const PcgState* gate = nullptr) {
  const mgpbd_ctx::Factored& F = ctx->fac;
  const int n = ctx->L[0].n;
  hipStream_t s = ctx->stream;
  if (xd) {
    fac_ogather<double>(ctx, xd, done, gate);
    with_rows(F, [&](auto rows) {
      k_fac_oapply<double, 1, decltype(rows)><<<gs, kBlock, 0, s>>>(n, rows, F.gradC, F.alpha,
                                                                    F.u64, xd, b, y, ctx->ws,
                                                                    done, gate);
    });
    return;
  }
  fac_ogather<float>(ctx, xf, done, gate);
  with_rows(F, [&](auto rows) {
    using RS = decltype(rows);
    if (mode == 0)
      k_fac_oapply<float, 0, RS><<<gs, kBlock, 0, s>>>(n, rows, F.gradC, F.alpha, F.u64, xf, b,
                                                       y, ctx->ws, done, gate);
    else if (mode == 1)
      k_fac_oapply<float, 1, RS><<<gs, kBlock, 0, s>>>(n, rows, F.gradC, F.alpha, F.u64, xf, b,
                                                       y, ctx->ws, done, gate);
    else
      k_fac_oapply<float, 2, RS><<<gs, kBlock, 0, s>>>(n, rows, F.gradC, F.alpha, F.u64, xf, b,
                                                       y, ctx->ws, done, gate);
  });
}

// One PCG iteration (cons/mgpcg_gpu.py:289-327) as a launch sequence on ctx->stream.
// z = M^-1 r for the PCG (first sweep included), FP32 or scaled reduced-precision V-cycle;
// scaled_ready: the PCG update already wrote the scaled level 0's b^ and first sweep.
static float* pcg_precondition(mgpbd_ctx* ctx, float* r, float* xv, bool first,
                               bool scaled_ready = false) {
  MgLevel& L = ctx->L[0];
  const int* done = &ctx->st->done;
  if (half_mode(ctx)) {
    if (!scaled_ready)
      k_half_first<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(L.n, L.isd, r, L.wb(0), L.wb(1),
                                                             ctx->omega, wl(ctx, 0), done);
    return vcycle_h(ctx, 0, L.wb(0), L.wb(1), done, L.wb(3));
  }
  if (!first) k_blas1<<<nblocks(L.n), kBlock, 0, ctx->stream>>>(3, L.n, 0.f, nullptr, xv);
  return vcycle_from(ctx, 0, r, xv, done);
}

// One PCG iteration (cons/mgpcg_gpu.py:289-327) as a launch sequence on ctx->stream.
static void pcg_iteration(mgpbd_ctx* ctx, float* x, double* x64, float* r, float* p, float* Ap,
                          float* xv, float* zold, bool first, int gn, int gs, bool prof) {
  MgLevel& L = ctx->L[0];
  const int n = L.n;
  hipStream_t s = ctx->stream;
  int pe = prof ? prof_begin(ctx, MGPBD_PROF_SPMV_L0) : -1;
  const size_t lds = sizeof(float) * kWinSeg * L.win_segs;
  if (ctx->fac.on)
    fac_outer(ctx, 0, p, nullptr, nullptr, Ap, &ctx->st->done, gs);
  else if (L.woff && x64)
    k_pcg_spmv_dot_w<true><<<gs, kBlock, lds, s>>>(n, L.n_slices, L.slice_ptr, L.woff, L.val,
                                                   L.seg_ptr, L.segs, p, Ap, ctx->ws, ctx->st);
  else if (L.woff)
    k_pcg_spmv_dot_w<false><<<gs, kBlock, lds, s>>>(n, L.n_slices, L.slice_ptr, L.woff, L.val,
                                                    L.seg_ptr, L.segs, p, Ap, ctx->ws, ctx->st);
  else if (x64)
    k_pcg_spmv_dot<true><<<gs, kBlock, 0, s>>>(n, L.n_slices, L.slice_ptr, L.idx, L.val, p, Ap,
                                               ctx->ws, ctx->st);
  else
    k_pcg_spmv_dot<false><<<gs, kBlock, 0, s>>>(n, L.n_slices, L.slice_ptr, L.idx, L.val, p, Ap,
                                                ctx->ws, ctx->st);
  prof_end(ctx, pe);
  k_pcg_fin<FIN_ALPHA><<<1, kBlock, 0, s>>>(ctx->ws, gs, ctx->st);
  // the update also writes the V-cycle's first sweep: from 1/diag on an FP32 level 0's first
  // iteration, from isd (with b^ = isd r, k_half_first's values) on a scaled level 0
  const bool fuse_first = first && !half_mode(ctx);
  const bool fuse_half = half_mode(ctx) && ctx->fused_launch;
  k_pcg_update<<<gn, kBlock, 0, s>>>(n, x, r, p, Ap, fuse_half ? L.isd : L.dinv,
                                     fuse_half ? L.wb(1) : (fuse_first ? xv : nullptr),
                                     ctx->omega, wl(ctx, 0), ctx->ws, ctx->st, x64,
                                     fuse_half ? L.wb(0) : nullptr);
  k_pcg_fin<FIN_RR><<<1, kBlock, 0, s>>>(ctx->ws, gn, ctx->st);
  float* zz = pcg_precondition(ctx, r, xv, first, fuse_half);
  if (ctx->flexible) {
    k_pcg_rz2<<<gn, kBlock, 0, s>>>(n, r, zz, zold, ctx->ws, ctx->st);
    k_pcg_fin_flex<<<1, kBlock, 0, s>>>(ctx->ws, gn, ctx->st);
    k_pcg_p_flex<<<nblocks(n), kBlock, 0, s>>>(n, zz, p, zold, ctx->st, 0);
  } else {
    k_pcg_rz<<<gn, kBlock, 0, s>>>(n, r, zz, ctx->ws, ctx->st);
    k_pcg_fin<FIN_BETA><<<1, kBlock, 0, s>>>(ctx->ws, gn, ctx->st);
    k_pcg_p<1><<<nblocks(n), kBlock, 0, s>>>(n, zz, p, ctx->st);
  }
}

int mgpbd_ctx_use_graphs(mgpbd_ctx* ctx, int enable, int poll) {
  CTX_CHECK("mgpbd_ctx_use_graphs");
  if (poll < 1 || poll > 64) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_ctx_use_graphs: poll in [1, 64]");
  ctx->use_graph = enable != 0;
  ctx->poll = poll;
  return 0;
}

static int pcg_solve(mgpbd_ctx* ctx, const float* b, float* x, const float* x_alt, bool guess,
                     float* work, int maxiter, double tol, int relative, int* iters_out,
                     double* info_out, int n_info) {
  if (!b || !x || !work || maxiter < 0 || (x_alt && (!guess || x_alt == x)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pcg_solve: invalid arguments");
  MgLevel& L = ctx->L[0];
  const int n = L.n;
  if (nblocks(64LL * L.n_slices) > kPartialStride)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pcg_solve: level 0 exceeds the reduction capacity");
  float* r = work;
  float* zold = work + (size_t)n;   // previous z (flexible PCG)
  float* p = work + 2 * (size_t)n;
  float* Ap = work + 3 * (size_t)n;
  // FP64 iterate (work + 4n floats, 8-byte aligned when work is)
  double* x64 = ctx->fp64_iterate ? reinterpret_cast<double*>(work + 4 * (size_t)n) : nullptr;
  float* xv = L.wb(1);              // level-0 x buffer receiving the first sweep
  const bool first = (ctx->n_levels == 1) || ctx->n_smooth > 0;
  const int* done = &ctx->st->done;
  const int gn = red_blocks(n);
  const int gs = nblocks(64LL * L.n_slices);
  // Graph replay runs on the library's own stream, ordered after the caller's stream; the
  // profiling mode launches eagerly on the caller's stream (events cannot be captured).
  const bool graph_mode = ctx->use_graph && !ctx->prof;
  hipStream_t caller = ctx->stream;
  if (graph_mode) {
    MGPBD_HIP_TRY(hipEventRecord(ctx->ev_in, caller));
    MGPBD_HIP_TRY(hipStreamWaitEvent(ctx->work, ctx->ev_in, 0));
    ctx->stream = ctx->work;
  }
  struct Restore {
    mgpbd_ctx* c; hipStream_t s;
    ~Restore() { c->stream = s; }
  } restore{ctx, caller};
  hipStream_t s = ctx->stream;

  // Replay lengths: a round's first replay runs as many iterations as the previous solve
  // took, the following ones single iterations up to one past that count, then `poll` at a
  // time. Iterations past convergence are skipped on the device's done flag and change
  // nothing, so the schedule only trades host polls against launches that do no work.
  const int first_len = std::max(1, std::min(ctx->pred_iters > 0 ? ctx->pred_iters : ctx->poll,
                                             (int)mgpbd_ctx::kGraphMax));
  const int pred_end = ctx->pred_iters > 0 ? ctx->pred_iters + 1 : 0;
  // with the factored level 0 every captured replay ends with the FP64 check, gated on the
  // device's done flag (a converged solve needs no second launch sequence and host wait)
  const bool tail_verify = graph_mode && ctx->verify && ctx->fac.on;
  // capture (or reuse) the graphs of those lengths before any work is queued on the stream
  if (graph_mode) {
    const void* key[4] = {b, x, work, L.work};
    if (std::memcmp(key, ctx->graph_key, sizeof(key)) != 0) {
      drop_graphs(ctx);
      std::memcpy(ctx->graph_key, key, sizeof(key));
    }
    const int lens[3] = {first_len, 1, ctx->poll};
    for (int len : lens) {
      if (ctx->graphs[len]) continue;
      hipGraph_t g = nullptr;
      MGPBD_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int c = 0; c < len; ++c) pcg_iteration(ctx, x, x64, r, p, Ap, xv, zold, first, gn, gs, false);
      if (tail_verify) {   // runs only once converged: the host reads rr_true with done
        if (x64) k_x_from64_gated<<<nblocks(n), kBlock, 0, s>>>(n, x64, x, ctx->st);
        fac_outer(ctx, 2, x, nullptr, b, nullptr, nullptr, gs, ctx->st);
        k_fin_true_gated<<<1, kBlock, 0, s>>>(ctx->ws, gs, ctx->st);
      }
      (void)hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, s);
      hipError_t ce = hipStreamEndCapture(s, &g);
      if (ce != hipSuccess) return mgpbd_fail(MGPBD_EHIP, std::string("graph capture: ") + hipGetErrorString(ce));
      ce = hipGraphInstantiate(&ctx->graphs[len], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ce != hipSuccess) {
        ctx->graphs[len] = nullptr;
        return mgpbd_fail(MGPBD_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ce));
      }
    }
  }

  const bool fuse_first = first && !half_mode(ctx);
  int issued = 0, total_iters = 0, restarts = 0;
  bool rescued = false, lag_rescued = false, first_round = true;
  const bool lagged_at_start = ctx->coarse_lagged;
  const int lag_limit = ctx->lag_limit < 0 ? std::max(8, 2 * ctx->fresh_iters + 2) : ctx->lag_limit;
  bool from_guess = guess;
  const float* alt = x_alt;
  double first_info[4] = {0.0, 0.0, 0.0, 0.0};   // rr0, start, rr_guess, rr_alt of round 0
  double rr_true = -1.0;
  bool nonfinite = false;   // NaN/Inf in ||b||^2 or the recurrence residual at a host poll
  for (;;) {
    k_pcg_reset<<<1, 1, 0, s>>>(ctx->st, tol, relative);
    if (from_guess) {
      // r = b - A x, q = b - A x_alt (q lives in Ap until the first iteration), then the
      // start with the smallest residual; a restart of the FP64 iterate takes r = b - A x64
      if (ctx->fac.on) {
        if (x64 && (restarts > 0 || rescued || lag_rescued))
          fac_outer(ctx, 1, nullptr, x64, b, r, nullptr, gs);
        else fac_outer(ctx, 1, x, nullptr, b, r, nullptr, gs);
        if (alt) fac_outer(ctx, 1, alt, nullptr, b, Ap, nullptr, gs);
      } else {
        if (x64 && (restarts > 0 || rescued || lag_rescued))
          k_resid64<<<gs, kBlock, 0, s>>>(n, L.n_slices, L.slice_ptr, L.idx, L.val, x64, b, r);
        else
          sell_launch<2>(ctx, L, x, r, nullptr, b, 0.f, nullptr, nullptr);
        if (alt)
          sell_launch<2>(ctx, L, alt, Ap, nullptr, b, 0.f, nullptr, nullptr);
      }
      k_pcg_guess_norms<<<gn, kBlock, 0, s>>>(n, b, r, alt ? Ap : nullptr, x, alt,
                                               ctx->start_energy ? 1 : 0, ctx->ws);
      k_pcg_fin_guess<<<1, kBlock, 0, s>>>(ctx->ws, gn, alt != nullptr,
                                           ctx->start_energy ? 1 : 0, ctx->st);
      k_pcg_select<<<nblocks(n), kBlock, 0, s>>>(n, b, x, r, alt, Ap, L.dinv,
                                                 fuse_first ? xv : nullptr, ctx->omega,
                                                 wl(ctx, 0), ctx->st);
      if (x64)
        k_x64_from_start<<<nblocks(n), kBlock, 0, s>>>(n, x, x64,
                                                       restarts > 0 || rescued || lag_rescued,
                                                       ctx->st);
    } else {
      k_pcg_init<<<gn, kBlock, 0, s>>>(n, b, x, r, L.dinv, fuse_first ? xv : nullptr,
                                        ctx->omega, wl(ctx, 0), ctx->ws, x64);
      k_pcg_fin<FIN_BB><<<1, kBlock, 0, s>>>(ctx->ws, gn, ctx->st);
    }
    float* zr = pcg_precondition(ctx, r, xv, first);
    k_pcg_rz<<<gn, kBlock, 0, s>>>(n, r, zr, ctx->ws, ctx->st);
    k_pcg_fin<FIN_RZ0><<<1, kBlock, 0, s>>>(ctx->ws, gn, ctx->st);
    if (ctx->flexible) k_pcg_p_flex<<<nblocks(n), kBlock, 0, s>>>(n, zr, p, zold, ctx->st, 1);
    else k_pcg_p<0><<<nblocks(n), kBlock, 0, s>>>(n, zr, p, ctx->st);
    MGPBD_LAUNCH_CHECK("mgpbd_pcg_solve(init)");

    int round_issued = 0;
    const bool watch = ctx->rescue && ctx->spectral && !ctx->fixed_weights && !rescued;
    double rr_poll[64];    // recurrence residual and iteration count at each host poll
    int it_poll[64];
    int n_poll = 0;
    bool stalled = false, lagging = false;
    const bool lag_watch = ctx->coarse_lagged && !lag_rescued && lag_limit > 0;
    while (issued < maxiter) {
      const int want = std::min(round_issued == 0 ? first_len
                                : round_issued < pred_end ? 1 : ctx->poll, maxiter - issued);
      if (graph_mode && ctx->graphs[want]) {
        MGPBD_HIP_TRY(hipGraphLaunch(ctx->graphs[want], s));
        issued += want;
        round_issued += want;
      } else {
        // profiling times every launch: one iteration per host check, so no launch past
        // convergence (skipped by the done flag) dilutes the per-kernel averages
        int chunk = ctx->prof ? 1 : want;
        for (int c = 0; c < chunk; ++c) pcg_iteration(ctx, x, x64, r, p, Ap, xv, zold, first, gn, gs, true);
        MGPBD_LAUNCH_CHECK("mgpbd_pcg_solve(iteration)");
        issued += chunk;
        round_issued += chunk;
        MGPBD_HIP_TRY(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
      }
      MGPBD_HIP_TRY(hipStreamSynchronize(s));
      if (int e = prof_flush(ctx)) return e;
      if (ctx->st_host->done) break;
      if (!std::isfinite(ctx->st_host->bb) || !std::isfinite(ctx->st_host->rr)) {
        nonfinite = true;   // NaN never meets the stop tests: stop polling at once
        break;
      }
      if (lag_watch && round_issued >= lag_limit) {
        lagging = true;   // levels 1.. older than level 0 and the solve runs long
        break;
      }
      if (watch) {
        rr_poll[n_poll % 64] = ctx->st_host->rr;
        it_poll[n_poll % 64] = round_issued;
        ++n_poll;
        // the latest poll at least 8 iterations back (within the kept history)
        int q = -1;
        for (int k = n_poll - 2; k >= 0 && k >= n_poll - 64; --k)
          if (it_poll[k % 64] <= round_issued - 8) { q = k; break; }
        if (round_issued >= 16 && q >= 0 &&
            ctx->st_host->rr > ctx->stall_ratio * rr_poll[q % 64]) {
          stalled = true;
          break;
        }
      }
    }
    if (round_issued == 0) {
      MGPBD_HIP_TRY(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
      MGPBD_HIP_TRY(hipStreamSynchronize(s));
    }
    if (nonfinite || !std::isfinite(ctx->st_host->bb)) {
      nonfinite = true;
      break;
    }
    const bool tail_verified = tail_verify && ctx->st_host->verified == 1;
    if (x64 && !tail_verified) {
      k_x_from64<<<nblocks(n), kBlock, 0, s>>>(n, x64, x);
      MGPBD_LAUNCH_CHECK("mgpbd_pcg_solve(x)");
    }
    if (first_round) {
      first_round = false;
      const PcgState& h = *ctx->st_host;
      first_info[0] = guess ? h.rr0 : h.bb;
      first_info[1] = (double)h.start;
      first_info[2] = h.rr_guess;
      first_info[3] = h.rr_alt;
    }
    total_iters += (stalled || lagging) ? round_issued : ctx->st_host->iters;
    if (!stalled && !lagging && restarts == 0 && !rescued && !lag_rescued) {
      ctx->pred_iters = std::max(1, ctx->st_host->iters);
      if (!lagged_at_start) ctx->fresh_iters = ctx->st_host->iters;
    }
    if (lagging && issued < maxiter) {
      // the lagged hierarchy no longer fits level 0: rebuild levels 1.. and resume
      lag_rescued = true;
      ++ctx->n_lag_rescues;
      if (int e = rebuild_coarse(ctx, true, 1)) return e;
      from_guess = true;
      alt = nullptr;
      continue;
    }
    if (stalled && issued < maxiter) {
      // the V-cycle no longer contracts: the spectral estimates lag the matrix
      rescued = true;
      ++ctx->n_rescues;
      if (int e = smoother_refresh(ctx, true, ctx->rescue_boost)) return e;
      from_guess = true;
      alt = nullptr;
      continue;
    }
    if (!ctx->verify) break;
    // FP64 residual of the returned x (every solve); a restart from x when a converged one
    // misses the tolerance
    if (!tail_verified) {
      if (ctx->fac.on)
        fac_outer(ctx, 2, x, nullptr, b, nullptr, nullptr, gs);
      else if (L.woff)
        k_true_resid_w<<<gs, kBlock, sizeof(float) * kWinSeg * L.win_segs, s>>>(
            n, L.n_slices, L.slice_ptr, L.woff, L.val, L.seg_ptr, L.segs, x, b, ctx->ws);
      else
        k_true_resid<<<gs, kBlock, 0, s>>>(n, L.n_slices, L.slice_ptr, L.idx, L.val, x, b, ctx->ws);
      k_fin_true<<<1, kBlock, 0, s>>>(ctx->ws, gs, ctx->st);
      MGPBD_LAUNCH_CHECK("mgpbd_pcg_solve(verify)");
      MGPBD_HIP_TRY(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
      MGPBD_HIP_TRY(hipStreamSynchronize(s));
    }
    rr_true = ctx->st_host->rr_true;
    if (ctx->verify == 2 || ctx->st_host->done != 1 || rr_true < ctx->st_host->tol2 ||
        issued >= maxiter || restarts >= 3)
      break;
    ++restarts;
    ++ctx->n_restarts;
    from_guess = true;
    alt = nullptr;
  }
  if (graph_mode) {
    MGPBD_HIP_TRY(hipEventRecord(ctx->ev_out, s));
    MGPBD_HIP_TRY(hipStreamWaitEvent(caller, ctx->ev_out, 0));
  }
  if (nonfinite) {
    const PcgState& h = *ctx->st_host;
    ctx->pred_iters = 0;
    return mgpbd_fail(MGPBD_ENONFINITE,
                      "mgpbd_pcg_solve: non-finite value (||b||^2 = " + std::to_string(h.bb) +
                          ", ||r||^2 = " + std::to_string(h.rr) + " after " +
                          std::to_string(issued) + " iterations)");
  }
  if (iters_out) *iters_out = total_iters;
  if (info_out) {
    const PcgState& h = *ctx->st_host;
    double reason = (double)h.done;
    if (ctx->verify == 1 && h.done == 1 && !(rr_true < h.tol2)) reason = 0.0;   // not verified
    const double v4[4] = {h.rr, h.bb, reason, rr_true};
    const double v8[8] = {h.rr, h.bb, reason, first_info[0], first_info[1], first_info[2],
                          first_info[3], rr_true};
    for (int i = 0; i < n_info; ++i) info_out[i] = n_info == 4 ? v4[i] : v8[i];
  }
  return 0;
}

int mgpbd_pcg_set_rescue(mgpbd_ctx* ctx, int enable, float boost, double stall_ratio) {
  CTX_CHECK("mgpbd_pcg_set_rescue");
  if (!(boost >= 1.0f && boost <= 4.0f) || !(stall_ratio >= 0.0 && stall_ratio <= 1.0))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pcg_set_rescue: boost in [1, 4], stall_ratio in [0, 1]");
  ctx->rescue = enable != 0;
  ctx->rescue_boost = boost;
  ctx->stall_ratio = stall_ratio;
  return 0;
}

int mgpbd_pcg_counters(mgpbd_ctx* ctx, int64_t* rescues, int64_t* restarts) {
  CTX_CHECK("mgpbd_pcg_counters");
  if (rescues) *rescues = ctx->n_rescues + ctx->n_lag_rescues;
  if (restarts) *restarts = ctx->n_restarts;
  return 0;
}

int mgpbd_pcg_set_lag_rescue(mgpbd_ctx* ctx, int limit) {
  CTX_CHECK("mgpbd_pcg_set_lag_rescue");
  if (limit < -1) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pcg_set_lag_rescue: limit >= -1");
  ctx->lag_limit = limit;
  return 0;
}

int mgpbd_pcg_set_start_criterion(mgpbd_ctx* ctx, int energy) {
  CTX_CHECK("mgpbd_pcg_set_start_criterion");
  ctx->start_energy = energy != 0;
  return 0;
}

int mgpbd_pcg_lag_rescues(mgpbd_ctx* ctx, int64_t* lag_rescues) {
  CTX_CHECK("mgpbd_pcg_lag_rescues");
  if (lag_rescues) *lag_rescues = ctx->n_lag_rescues;
  return 0;
}

int mgpbd_pcg_set_fp64_iterate(mgpbd_ctx* ctx, int enable) {
  CTX_CHECK("mgpbd_pcg_set_fp64_iterate");
  const bool on = enable != 0;
  if (on != ctx->fp64_iterate) drop_graphs(ctx);   // the captured iteration changes
  ctx->fp64_iterate = on;
  return 0;
}

int mgpbd_pcg_set_verify(mgpbd_ctx* ctx, int enable) {
  CTX_CHECK("mgpbd_pcg_set_verify");
  if (enable < 0 || enable > 2) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_pcg_set_verify: 0, 1 or 2");
  if ((enable != 0) != (ctx->verify != 0)) drop_graphs(ctx);   // the replays' check tail
  ctx->verify = enable;
  return 0;
}

int mgpbd_pcg_solve(mgpbd_ctx* ctx, const float* b, float* x, float* work, int maxiter,
                    double tol, int relative, int* iters_out, double* info_out) {
  CTX_CHECK("mgpbd_pcg_solve");
  if (int e = check_ready(ctx)) return e;
  return pcg_solve(ctx, b, x, nullptr, false, work, maxiter, tol, relative, iters_out,
                   info_out, 4);
}

int mgpbd_pcg_solve_guess(mgpbd_ctx* ctx, const float* b, float* x, const float* x_alt,
                          float* work, int maxiter, double tol, int relative, int* iters_out,
                          double* info_out) {
  CTX_CHECK("mgpbd_pcg_solve_guess");
  if (int e = check_ready(ctx)) return e;
  return pcg_solve(ctx, b, x, x_alt, true, work, maxiter, tol, relative, iters_out, info_out,
                   8);
}

void* mgpbd_internal_stream(mgpbd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mgpbd_constraints_noalloc(void* stream, int n_tets, const int32_t* t_i, const float* v_p,
                              const float* Dm_inv, const float* alpha, const float* lambda,
                              float* C, float* gradC, float* rhs, double* norm2) {
  if (n_tets < 0 || n_tets > kNoAllocMaxTets)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints_noalloc: n_tets must be in [0, 2^28] "
                                    "(got " + std::to_string(n_tets) + ")");
  if (n_tets > 0 && (!t_i || !v_p || !Dm_inv || !alpha || !lambda || !C || !gradC || !rhs))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints_noalloc: null array");
  if (!norm2) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_constraints_noalloc: null norm2");
  double* part = nullptr;
  if (int e = scratch_address(&part)) return e;
  const int g = red_blocks(n_tets);
  if (g > kNoAllocPartials)
    return mgpbd_fail(MGPBD_EINTERNAL, "mgpbd_constraints_noalloc: grid exceeds the scratch");
  // every block stores its partial and the finaliser sums exactly g of them: nothing in the
  // scratch has to start at zero
  RedWS ws{part};
  hipStream_t s = (hipStream_t)stream;
  k_constraints<false><<<g, kBlock, 0, s>>>(n_tets, reinterpret_cast<const int4*>(t_i), v_p, Dm_inv,
                                     alpha, lambda, C, gradC, rhs, ws);
  k_fin_sum<<<1, kBlock, 0, s>>>(ws, g, 0, 1, norm2);
  MGPBD_LAUNCH_CHECK("mgpbd_constraints_noalloc");
  return 0;
}

}  // extern "C"
