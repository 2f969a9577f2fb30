// 3x3 rotation-variant SVD, F = U diag(s) V^T with U, V proper rotations and the sign of
// det(F) carried by the smallest singular value -- the convention of Taichi's ti.svd, which
// the reference calls in cons/mgpbd_gpu.py:236,372 (McAdams et al. 2011 there).
//
// Method: one-sided (Hestenes) Jacobi. Columns of B = F V are orthogonalised by plane
// rotations accumulated into V; then s_i = |b_i|, u_i = b_i / s_i and signs are fixed so that
// det V = det U = +1. The solver runs it in FP64 (per tet ~1 kFLOP, far below the kernel's
// memory time): dC/dF = sum_i (s_i - 1)/C u_i v_i^T divides by C, so FP32 rounding in s
// would turn into a direction error of ~eps/C for slightly strained tets. The quantities the
// solver uses do not depend on the order of the singular values.
#pragma once

#include <hip/hip_runtime.h>

namespace mgpbd {

template <typename T> struct SvdEps;
template <> struct SvdEps<float> { static constexpr float rel = 1e-7f, tiny = 1e-30f; };
template <> struct SvdEps<double> { static constexpr double rel = 1e-15, tiny = 1e-300; };

__device__ __forceinline__ float dot3(const float* a, const float* b) {
  return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}

// Columns stored as col[c][r]. Returns false when the pair is already orthogonal.
template <typename T>
__device__ __forceinline__ bool jacobi_rotate(T (&B)[3][3], T (&V)[3][3], int p, int q) {
  T a = dot3(B[p], B[p]);
  T b = dot3(B[q], B[q]);
  T g = dot3(B[p], B[q]);
  if (fabs(g) <= SvdEps<T>::tiny || fabs(g) <= SvdEps<T>::rel * sqrt(a * b)) return false;
  T zeta = (b - a) / (T(2) * g);
  T t = T(1) / (fabs(zeta) + sqrt(zeta * zeta + T(1)));
  t = zeta >= T(0) ? t : -t;
  T c = T(1) / sqrt(t * t + T(1));
  T s = c * t;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    T bp = B[p][r], bq = B[q][r];
    B[p][r] = c * bp - s * bq;
    B[q][r] = s * bp + c * bq;
    T vp = V[p][r], vq = V[q][r];
    V[p][r] = c * vp - s * vq;
    V[q][r] = s * vp + c * vq;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ __forceinline__ T det_cols(const T (&M)[3][3]) {
  T c[3];
  cross3(M[1], M[2], c);
  return dot3(M[0], c);
}

// F row-major in, U and V as columns (U[c][r]), s signed.
template <typename T>
__device__ __forceinline__ void svd3(const T F[9], T (&U)[3][3], T (&s)[3], T (&V)[3][3]) {
  T B[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      B[c][r] = F[3 * r + c];
      V[c][r] = (r == c) ? T(1) : T(0);
    }
  for (int sweep = 0; sweep < 12; ++sweep) {
    bool any = jacobi_rotate(B, V, 0, 1);
    any |= jacobi_rotate(B, V, 0, 2);
    any |= jacobi_rotate(B, V, 1, 2);
    if (!any) break;
  }
  T sg[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) sg[c] = sqrt(dot3(B[c], B[c]));
  int m = 0;
  if (sg[1] < sg[m]) m = 1;
  if (sg[2] < sg[m]) m = 2;
  int a = (m + 1) % 3, b = (m + 2) % 3;
  // make V a rotation by flipping the pair (v_m, b_m); F = sum b_i v_i^T is unchanged
  if (det_cols(V) < T(0)) {
#pragma unroll
    for (int r = 0; r < 3; ++r) { V[m][r] = -V[m][r]; B[m][r] = -B[m][r]; }
  }
  const T tiny = SvdEps<T>::tiny;
  // u_a, u_b: normalised columns (with a fallback basis for collapsed tets)
  if (sg[a] > tiny) {
    T inv = T(1) / sg[a];
#pragma unroll
    for (int r = 0; r < 3; ++r) U[a][r] = B[a][r] * inv;
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r) U[a][r] = (r == a) ? T(1) : T(0);
  }
  if (sg[b] > tiny) {
    T inv = T(1) / sg[b];
#pragma unroll
    for (int r = 0; r < 3; ++r) U[b][r] = B[b][r] * inv;
  } else {
    // any unit vector orthogonal to u_a
    T e[3] = {T(0), T(0), T(0)};
    int k = fabs(U[a][0]) < T(0.57) ? 0 : (fabs(U[a][1]) < T(0.57) ? 1 : 2);
    e[k] = T(1);
    T d = dot3(e, U[a]);
    T w[3] = {e[0] - d * U[a][0], e[1] - d * U[a][1], e[2] - d * U[a][2]};
    T inv = T(1) / sqrt(dot3(w, w));
#pragma unroll
    for (int r = 0; r < 3; ++r) U[b][r] = w[r] * inv;
  }
  T sm = sg[m];
  if (sm > tiny) {
    T inv = T(1) / sm;
#pragma unroll
    for (int r = 0; r < 3; ++r) U[m][r] = B[m][r] * inv;
    if (det_cols(U) < T(0)) {
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
