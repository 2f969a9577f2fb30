// 3x3 rotation-variant SVD, F = U diag(s) V^T with U, V proper rotations and the sign of
// det(F) carried by the smallest singular value -- the convention of Taichi's ti.svd, which
// the reference calls in cons/mgpbd_gpu.py:236,372 (McAdams et al. 2011 there).
//
// Method: one-sided (Hestenes) Jacobi. Columns of B = F V are orthogonalised by plane
// rotations accumulated into V; then s_i = |b_i|, u_i = b_i / s_i and signs are fixed so that
// det V = det U = +1. Works in FP32 on the GPU; the quantities the solver uses
// (C = |s - 1| and dC/dF = sum_i (s_i - 1)/C u_i v_i^T) do not depend on the order of the
// singular values.
#pragma once

#include <hip/hip_runtime.h>

namespace mgpbd {

struct Mat3 { float m[3][3]; };   // m[row][col]

__device__ __forceinline__ float dot3(const float* a, const float* b) {
  return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}

// Columns stored as col[c][r].
__device__ __forceinline__ void jacobi_rotate(float (&B)[3][3], float (&V)[3][3], int p, int q) {
  float a = dot3(B[p], B[p]);
  float b = dot3(B[q], B[q]);
  float g = dot3(B[p], B[q]);
  if (fabsf(g) <= 1e-30f || fabsf(g) <= 1e-9f * sqrtf(a * b)) return;
  float zeta = (b - a) / (2.0f * g);
  float t = 1.0f / (fabsf(zeta) + sqrtf(fmaf(zeta, zeta, 1.0f)));
  t = zeta >= 0.0f ? t : -t;
  float c = 1.0f / sqrtf(fmaf(t, t, 1.0f));
  float s = c * t;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float bp = B[p][r], bq = B[q][r];
    B[p][r] = c * bp - s * bq;
    B[q][r] = s * bp + c * bq;
    float vp = V[p][r], vq = V[q][r];
    V[p][r] = c * vp - s * vq;
    V[q][r] = s * vp + c * vq;
  }
}

__device__ __forceinline__ void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ float det_cols(const float (&M)[3][3]) {
  float c[3];
  cross3(M[1], M[2], c);
  return dot3(M[0], c);
}

// F row-major in, U and V as columns (U[c][r]), s signed.
__device__ __forceinline__ void svd3(const float F[9], float (&U)[3][3], float (&s)[3],
                                     float (&V)[3][3]) {
  float B[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      B[c][r] = F[3 * r + c];
      V[c][r] = (r == c) ? 1.0f : 0.0f;
    }
  for (int sweep = 0; sweep < 6; ++sweep) {
    jacobi_rotate(B, V, 0, 1);
    jacobi_rotate(B, V, 0, 2);
    jacobi_rotate(B, V, 1, 2);
  }
  float sg[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) sg[c] = sqrtf(dot3(B[c], B[c]));
  int m = 0;
  if (sg[1] < sg[m]) m = 1;
  if (sg[2] < sg[m]) m = 2;
  int a = (m + 1) % 3, b = (m + 2) % 3;
  // make V a rotation by flipping the pair (v_m, b_m); F = sum b_i v_i^T is unchanged
  if (det_cols(V) < 0.0f) {
#pragma unroll
    for (int r = 0; r < 3; ++r) { V[m][r] = -V[m][r]; B[m][r] = -B[m][r]; }
  }
  const float tiny = 1e-30f;
  // u_a, u_b: normalised columns (with a fallback basis for collapsed tets)
  if (sg[a] > tiny) {
    float inv = 1.0f / sg[a];
#pragma unroll
    for (int r = 0; r < 3; ++r) U[a][r] = B[a][r] * inv;
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r) U[a][r] = (r == a) ? 1.0f : 0.0f;
  }
  if (sg[b] > tiny) {
    float inv = 1.0f / sg[b];
#pragma unroll
    for (int r = 0; r < 3; ++r) U[b][r] = B[b][r] * inv;
  } else {
    // any unit vector orthogonal to u_a
    float e[3] = {0.f, 0.f, 0.f};
    int k = fabsf(U[a][0]) < 0.57f ? 0 : (fabsf(U[a][1]) < 0.57f ? 1 : 2);
    e[k] = 1.0f;
    float d = dot3(e, U[a]);
    float w[3] = {e[0] - d * U[a][0], e[1] - d * U[a][1], e[2] - d * U[a][2]};
    float inv = 1.0f / sqrtf(dot3(w, w));
#pragma unroll
    for (int r = 0; r < 3; ++r) U[b][r] = w[r] * inv;
  }
  float sm = sg[m];
  if (sm > tiny) {
    float inv = 1.0f / sm;
#pragma unroll
    for (int r = 0; r < 3; ++r) U[m][r] = B[m][r] * inv;
    if (det_cols(U) < 0.0f) {
#pragma unroll
      for (int r = 0; r < 3; ++r) U[m][r] = -U[m][r];
      sm = -sm;
    }
  } else {
    // u_m completes a right-handed basis: u_m = u_a x u_b
    cross3(U[a], U[b], U[m]);
  }
  s[0] = sg[0]; s[1] = sg[1]; s[2] = sg[2];
  s[m] = sm;
}

}  // namespace mgpbd
