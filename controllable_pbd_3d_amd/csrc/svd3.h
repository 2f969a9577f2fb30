// 3x3 rotation-variant SVD, F = U diag(s) V^T with U, V proper rotations and the sign of
// det(F) carried by the smallest singular value -- the convention of Taichi's ti.svd, which
// the reference calls in cons/mgpbd_gpu.py:236,372 (McAdams et al. 2011 there).
//
// Method: one-sided (Hestenes) Jacobi. Columns of B = F V are orthogonalised by plane
// rotations accumulated into V; then s_i = |b_i|, u_i = b_i / s_i and signs are fixed so that
// det V = det U = +1. The solver runs it in FP64: dC/dF = sum_i (s_i - 1)/C u_i v_i^T divides
// by C, so FP32 rounding in s would turn into a direction error of ~eps/C for slightly
// strained tets. The quantities the solver uses do not depend on the order of the singular
// values. The ARAP constraint needs only R = U V^T, which polar_rotation (below) finds for
// det F > 0 at about a fifth of the Jacobi sweeps' FP64 work; svd3 covers the other tets.
#pragma once

#include <hip/hip_runtime.h>

namespace mgpbd {

template <typename T> struct SvdEps;
template <> struct SvdEps<float> { static constexpr float rel = 1e-7f, tiny = 1e-30f; };
template <> struct SvdEps<double> { static constexpr double rel = 1e-15, tiny = 1e-300; };

__host__ __device__ __forceinline__ float dot3(const float* a, const float* b) {
  return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
__host__ __device__ __forceinline__ double dot3(const double* a, const double* b) {
  return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}

// Columns stored as col[c][r]. Returns false when the pair is already orthogonal.
template <typename T>
__host__ __device__ __forceinline__ bool jacobi_rotate(T (&B)[3][3], T (&V)[3][3], int p, int q) {
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
__host__ __device__ __forceinline__ void cross3(const T* a, const T* b, T* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__host__ __device__ __forceinline__ T det_cols(const T (&M)[3][3]) {
  T c[3];
  cross3(M[1], M[2], c);
  return dot3(M[0], c);
}

// F row-major in, U and V as columns (U[c][r]), s signed.
template <typename T>
__host__ __device__ __forceinline__ void svd3(const T F[9], T (&U)[3][3], T (&s)[3], T (&V)[3][3]) {
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

// Rotation R = U V^T of F's rotation-variant SVD without the SVD, for det F > 0: then R is
// F's orthogonal polar factor, the limit of Newton's iteration X <- (X + X^-T) / 2 from X = F
// (Higham 1986; X^-T = cof(X) / det X, quadratic convergence, no scaling). Stops once a step
// moves X by less than 1e-8 (Frobenius): the error left is then about the step squared, below
// FP64 rounding. Returns false -- the caller runs svd3 -- when det X is not safely positive
// (inverted or collapsed tets: their R is not the polar factor) or after kPolarIters steps
// (singular values far from 1; an interior strain of 10 % converges in 4 steps).
// Then F - R = U (S - I) V^T, so the ARAP constraint sqrt(sum (s_i - 1)^2) is ||F - R||_F and
// its gradient with respect to F is (F - R) / C.
constexpr int kPolarIters = 10;
template <typename T>
__host__ __device__ __forceinline__ bool polar_rotation(const T F[9], T (&R)[9]) {
  T X[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) X[k] = F[k];
  const T f2 = dot3(F, F) + dot3(F + 3, F + 3) + dot3(F + 6, F + 6);   // ||F||_F^2
  for (int it = 0; it < kPolarIters; ++it) {
    T C[9];
    C[0] = X[4] * X[8] - X[5] * X[7];
    C[1] = X[5] * X[6] - X[3] * X[8];
    C[2] = X[3] * X[7] - X[4] * X[6];
    C[3] = X[2] * X[7] - X[1] * X[8];
    C[4] = X[0] * X[8] - X[2] * X[6];
    C[5] = X[1] * X[6] - X[0] * X[7];
    C[6] = X[1] * X[5] - X[2] * X[4];
    C[7] = X[2] * X[3] - X[0] * X[5];
    C[8] = X[0] * X[4] - X[1] * X[3];
    const T det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
    // det F > 1e-3 (||F||_F^2 / 3)^{3/2}: positive and not nearly collapsed (the iterates
    // keep F's singular vectors and the sign of its determinant, and only grow better
    // conditioned)
    if (it == 0 && !(det > T(0) && det * det > T(1e-6) * (f2 * f2 * f2) / T(27))) return false;
    const T h = T(0.5) / det;
    T step = T(0);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const T nx = T(0.5) * X[k] + h * C[k];
      const T d = nx - X[k];
      step += d * d;
      X[k] = nx;
    }
    if (step <= T(1e-16)) {
#pragma unroll
      for (int k = 0; k < 9; ++k) R[k] = X[k];
      return true;
    }
  }
  return false;
}

}  // namespace mgpbd
