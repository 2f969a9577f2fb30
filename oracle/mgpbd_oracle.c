/*
 * ORACLE (test infrastructure only) -- plain-C CPU restatement of the numeric part of the
 * reference MGPBD path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker / CPU baseline; the product never does.
 *
 * Built twice by oracle/Makefile: liboracle_f32.so (REAL = float, FP32 arithmetic like the
 * reference's Taichi kernels) and liboracle_f64.so (REAL = double, the FP64 check).
 * Compiled with -ffp-contract=off: every fused multiply-add below is an explicit FMA(),
 * written in the same order as the GPU kernels, so the FP32 build reproduces their SpMV,
 * Jacobi, A fill, coarse refill, restriction, prolongation and scatter bit for bit on the
 * same inputs. Reductions are sequential FP64.
 *
 * Reference functions restated (file:line in JamesYang-7/Controllable_PBD_3D):
 *   orc_rest_shape      cons/mgpbd_gpu.py:136-151   (_precompute_rest_shape; FP64 inverse)
 *   orc_predict         cons/mgpbd_gpu.py:206-213   (make_prediction)
 *   orc_update_vel      cons/mgpbd_gpu.py:215-219
 *   orc_constraints     cons/mgpbd_gpu.py:225-267 + cons/mgpbd.py:25-68 (C, dC/dx, rhs, |rhs|)
 *                       with an FP64 SVD (Jacobi eigen-decomposition of F^T F)
 *   orc_fill_A          cons/mgpbd_gpu.py:273-309 + cons/mgpbd.py:71-84
 *   orc_cache_diag_inv  cons/sparse_gpu.py:107-121
 *   orc_spmv/orc_jacobi cons/sparse_gpu.py:123-144
 *   orc_refill          cons/mgpcg_gpu.py:183-198
 *   orc_vcycle          cons/mgpcg_gpu.py:214-259
 *   orc_pcg             cons/mgpcg_gpu.py:261-329
 *   orc_scatter         cons/mgpbd_gpu.py:330-354
 *   orc_dual_norm2      cons/mgpbd_gpu.py:360-377
 *   orc_apply_dpos      cons/mgpbd_gpu.py:345-349
 *
 * Volume constraint (extension; the reference MGPBD solve has one ARAP constraint per tet):
 *   orc_rest_shape_vol, orc_constraints_vol, orc_dual_norm2_vol -- constraint 2t is tet t's
 *   ARAP constraint as above, constraint 2t+1 its volume constraint C_H = det F - 1 with
 *   dC_H/dF = cof(F) (the hydrostatic constraint of cons/deform3d.py:81-91) and compliance
 *   1 / (lambda dt^2 vol), lambda the Lame parameter (cons/vbd.py:44-47).
 *
 * Threads: OpenMP over rows / tets / vertices. Every row-parallel loop computes each output
 * from its own inputs in the sequential order, so results do not depend on the thread count;
 * FP64 reductions sum fixed chunks of ORC_CHUNK terms in parallel and the chunk totals in
 * chunk order (the same total for any thread count).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_CHUNK 8192
#define ORC_PAR_MIN 4096   /* shorter loops stay on one thread */

#ifdef ORACLE_F64
typedef double real;
#define FMA(a, b, c) fma((a), (b), (c))
#define RABS(a) fabs(a)
#else
typedef float real;
#define FMA(a, b, c) fmaf((a), (b), (c))
#define RABS(a) fabsf(a)
#endif

/* Threads of the OpenMP loops (the CPU baseline times 1 thread and all cores). */
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- rest shape (FP64) */
/* Dm^-1 of tet t into Dm_inv, returns vol = |det Dm| / 6 (FP64). */
static double rest_one(int t, const int32_t* t_i, const real* xr, real* Dm_inv) {
  {
    const int32_t* v = t_i + 4 * (int64_t)t;
    double D[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) D[r][c] = (double)xr[3 * v[c + 1] + r] - (double)xr[3 * v[0] + r];
    double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    double c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    double c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
    double inv[9] = {
        c00 / det, (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det, (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det,
        c01 / det, (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det, (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det,
        c02 / det, (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det, (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det};
    for (int k = 0; k < 9; ++k) Dm_inv[9 * (int64_t)t + k] = (real)inv[k];
    return fabs(det) / 6.0;
  }
}

void orc_rest_shape(int nt, const int32_t* t_i, const real* xr, const real* mu_t, double mu,
                    double dt, real* Dm_inv, real* vol_out, real* alpha) {
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; ++t) {
    double vol = rest_one(t, t_i, xr, Dm_inv);
    vol_out[t] = (real)vol;
    double m = mu_t ? (double)mu_t[t] : mu;
    alpha[t] = (real)(1.0 / (m * dt * dt * vol));
  }
}

/* As orc_rest_shape with two constraints per tet: alpha[2t] = 1/(mu dt^2 vol) (ARAP),
 * alpha[2t+1] = 1/(lambda dt^2 vol) (volume); lam_t (per tet) overrides the scalar lam. */
void orc_rest_shape_vol(int nt, const int32_t* t_i, const real* xr, const real* mu_t, double mu,
                        const real* lam_t, double lam, double dt, real* Dm_inv, real* vol_out,
                        real* alpha) {
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; ++t) {
    double vol = rest_one(t, t_i, xr, Dm_inv);
    vol_out[t] = (real)vol;
    double m = mu_t ? (double)mu_t[t] : mu;
    double l = lam_t ? (double)lam_t[t] : lam;
    alpha[2 * (int64_t)t] = (real)(1.0 / (m * dt * dt * vol));
    alpha[2 * (int64_t)t + 1] = (real)(1.0 / (l * dt * dt * vol));
  }
}

/* ---------------------------------------------------------------- time integration */
void orc_predict(int nv, real* x, real* v, real* x_old, const real* invm, const real* fext,
                 real gx, real gy, real gz, real dt, real damp, int hold_pinned) {
  real g[3] = {gx, gy, gz};
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nv; ++k) {
    int pinned = hold_pinned && !(invm[k] > 0);
    for (int r = 0; r < 3; ++r) {
      int64_t q = 3 * (int64_t)k + r;
      if (pinned) { v[q] = 0; x_old[q] = x[q]; continue; }
      real vv = v[q] + dt * g[r];
      if (fext) vv = vv + dt * invm[k] * fext[q];
      vv = vv * damp;
      v[q] = vv;
      x_old[q] = x[q];
      x[q] = x[q] + dt * vv;
    }
  }
}

void orc_update_vel(int nv, const real* x, const real* x_old, real* v, real dt, real damp) {
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < 3 * (int64_t)nv; ++q) v[q] = damp * (x[q] - x_old[q]) / dt;
}

void orc_apply_dpos(int nv, real* x, const real* dpos, const real* invm, real omega) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nv; ++k)
    if (invm[k] > 0)
      for (int r = 0; r < 3; ++r) x[3 * (int64_t)k + r] = FMA(omega, dpos[3 * (int64_t)k + r], x[3 * (int64_t)k + r]);
}

/* ---------------------------------------------------------------- FP64 SVD */
/* Symmetric Jacobi on S = F^T F gives V; B = F V has orthogonal columns b_i = s_i u_i.
 * U, V are made proper rotations and det(F)'s sign is put on the smallest s_i. */
static void svd3_f64(const double F[3][3], double U[3][3], double s[3], double V[3][3]) {
  double S[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      S[i][j] = 0;
      for (int k = 0; k < 3; ++k) S[i][j] += F[k][i] * F[k][j];
      V[i][j] = (i == j);
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = fabs(S[0][1]) + fabs(S[0][2]) + fabs(S[1][2]);
    double scale = fabs(S[0][0]) + fabs(S[1][1]) + fabs(S[2][2]);
    if (off <= 1e-300 || off <= 1e-17 * scale) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (S[p][q] == 0.0) continue;
        double theta = (S[q][q] - S[p][p]) / (2.0 * S[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 3; ++k) {  /* S <- J^T S J */
          double skp = S[k][p], skq = S[k][q];
          S[k][p] = c * skp - sn * skq;
          S[k][q] = sn * skp + c * skq;
        }
        for (int k = 0; k < 3; ++k) {
          double spk = S[p][k], sqk = S[q][k];
          S[p][k] = c * spk - sn * sqk;
          S[q][k] = sn * spk + c * sqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - sn * vkq;
          V[k][q] = sn * vkp + c * vkq;
        }
      }
  }
  double B[3][3];   /* B = F V, columns */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      B[i][j] = 0;
      for (int k = 0; k < 3; ++k) B[i][j] += F[i][k] * V[k][j];
    }
  double sg[3];
  for (int j = 0; j < 3; ++j) sg[j] = sqrt(B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j]);
  int m = 0;
  if (sg[1] < sg[m]) m = 1;
  if (sg[2] < sg[m]) m = 2;
  int a = (m + 1) % 3, b = (m + 2) % 3;
  double detV = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) -
                V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
  if (detV < 0)
    for (int r = 0; r < 3; ++r) { V[r][m] = -V[r][m]; B[r][m] = -B[r][m]; }
  for (int r = 0; r < 3; ++r) {
    U[r][a] = sg[a] > 1e-300 ? B[r][a] / sg[a] : (r == a);
    U[r][b] = sg[b] > 1e-300 ? B[r][b] / sg[b] : (r == b);
  }
  double sm = sg[m];
  double cr[3] = {U[1][a] * U[2][b] - U[2][a] * U[1][b], U[2][a] * U[0][b] - U[0][a] * U[2][b],
                  U[0][a] * U[1][b] - U[1][a] * U[0][b]};
  if (sm > 1e-300) {
    for (int r = 0; r < 3; ++r) U[r][m] = B[r][m] / sm;
    if (cr[0] * U[0][m] + cr[1] * U[1][m] + cr[2] * U[2][m] < 0) {
      for (int r = 0; r < 3; ++r) U[r][m] = -U[r][m];
      sm = -sm;
    }
  } else {
    for (int r = 0; r < 3; ++r) U[r][m] = cr[r];
  }
  s[0] = sg[0]; s[1] = sg[1]; s[2] = sg[2];
  s[m] = sm;
}

/* F = Ds Dm^-1 of one tet from positions (double); B receives Dm^-1. */
static void def_grad(const double p[4][3], const real* Bm, double B[3][3], double F[3][3]) {
  double Ds[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      Ds[r][c] = p[c + 1][r] - p[0][r];
      B[r][c] = (double)Bm[3 * r + c];
    }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) F[r][c] = Ds[r][0] * B[0][c] + Ds[r][1] * B[1][c] + Ds[r][2] * B[2][c];
}

/* vertex gradients from dC/dF = G: dC/dDs = G B^T, column k-1 is vertex k's gradient
 * (cons/mgpbd.py:41-67), vertex 0 takes minus their sum */
static void grad_from_dF(const double G[3][3], const double B[3][3], double g[4][3]) {
  for (int k = 1; k <= 3; ++k)
    for (int r = 0; r < 3; ++r) g[k][r] = G[r][0] * B[k - 1][0] + G[r][1] * B[k - 1][1] + G[r][2] * B[k - 1][2];
  for (int r = 0; r < 3; ++r) g[0][r] = -g[1][r] - g[2][r] - g[3][r];
}

/* ARAP C = ||sigma - 1|| and dC/dx of one tet from positions (double). g[4][3]; returns C. */
static double tet_C_grad(const double p[4][3], const real* Bm, double g[4][3], int want_grad) {
  double B[3][3], F[3][3];
  def_grad(p, Bm, B, F);
  double U[3][3], s[3], V[3][3];
  svd3_f64(F, U, s, V);
  double C = sqrt((s[0] - 1) * (s[0] - 1) + (s[1] - 1) * (s[1] - 1) + (s[2] - 1) * (s[2] - 1));
  if (!want_grad) return C;
  double d[3] = {(s[0] - 1) / C, (s[1] - 1) / C, (s[2] - 1) / C};
  double G[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) G[r][c] = U[r][0] * V[c][0] * d[0] + U[r][1] * V[c][1] * d[1] + U[r][2] * V[c][2] * d[2];
  grad_from_dF(G, B, g);
  return C;
}

/* Volume constraint C_H = det F - 1 of one tet, dC_H/dF = cof(F): column c is the cross
 * product of F's other two columns in cyclic order (cons/deform3d.py:81-91). */
static double tet_vol_grad(const double p[4][3], const real* Bm, double g[4][3], int want_grad) {
  double B[3][3], F[3][3];
  def_grad(p, Bm, B, F);
  double G[3][3];
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    G[0][c] = F[1][a] * F[2][b] - F[2][a] * F[1][b];
    G[1][c] = F[2][a] * F[0][b] - F[0][a] * F[2][b];
    G[2][c] = F[0][a] * F[1][b] - F[1][a] * F[0][b];
  }
  double det = F[0][0] * G[0][0] + F[1][0] * G[1][0] + F[2][0] * G[2][0];
  if (want_grad) grad_from_dF(G, B, g);
  return det - 1.0;
}

static void tet_positions(const int32_t* t_i, int t, const real* x, const real* dx, real omega,
                          double p[4][3]) {
  for (int l = 0; l < 4; ++l)
    for (int r = 0; r < 3; ++r) {
      int64_t q = 3 * (int64_t)t_i[4 * (int64_t)t + l] + r;
      real xv = x[q];
      if (dx) xv = xv + omega * dx[q];
      p[l][r] = (double)xv;
    }
}

/* Deterministic parallel FP64 sum: fn(lo, hi, arg) returns the sequential sum over rows
 * [lo, hi); chunks of ORC_CHUNK rows are summed in parallel, their totals in chunk order. */
typedef double (*orc_range_fn)(int lo, int hi, void* arg);
static double chunked_sum(int n, orc_range_fn fn, void* arg) {
  int nch = (n + ORC_CHUNK - 1) / ORC_CHUNK;
  if (nch <= 1) return n > 0 ? fn(0, n, arg) : 0.0;
  double* part = malloc(sizeof(double) * (size_t)nch);
#pragma omp parallel for schedule(static)
  for (int ch = 0; ch < nch; ++ch) {
    int lo = ch * ORC_CHUNK, hi = lo + ORC_CHUNK < n ? lo + ORC_CHUNK : n;
    part[ch] = fn(lo, hi, arg);
  }
  double tot = 0.0;
  for (int ch = 0; ch < nch; ++ch) tot += part[ch];
  free(part);
  return tot;
}

typedef struct {
  const int32_t* t_i; const real *x, *dx; real omega; const real *Dm_inv, *alpha, *lambda;
  real *C_out, *gradC, *rhs;
  int volume;   /* 0: one ARAP constraint per tet; 1: ARAP (2t) and volume (2t+1) */
} ConsArgs;

static void store_cons(const ConsArgs* a, int64_t c, double C, const double g[4][3],
                       int zero_small, double* acc) {
  real Cr = (real)C;
  a->C_out[c] = Cr;
  if (!zero_small || Cr > (real)1e-6) {
    for (int l = 0; l < 4; ++l)
      for (int r = 0; r < 3; ++r) a->gradC[12 * c + 3 * l + r] = (real)g[l][r];
  } else {
    for (int k = 0; k < 12; ++k) a->gradC[12 * c + k] = 0;
  }
  real b = -(Cr + a->alpha[c] * a->lambda[c]);
  a->rhs[c] = b;
  *acc += (double)b * (double)b;
}

static double cons_range(int lo, int hi, void* arg) {
  const ConsArgs* a = (const ConsArgs*)arg;
  double acc = 0.0;
  for (int t = lo; t < hi; ++t) {
    double p[4][3], g[4][3];
    tet_positions(a->t_i, t, a->x, NULL, 0, p);
    const int64_t c0 = a->volume ? 2 * (int64_t)t : t;
    double C = tet_C_grad(p, a->Dm_inv + 9 * (int64_t)t, g, 1);
    store_cons(a, c0, C, g, 1, &acc);
    if (a->volume) {
      double Ch = tet_vol_grad(p, a->Dm_inv + 9 * (int64_t)t, g, 1);
      store_cons(a, c0 + 1, Ch, g, 0, &acc);
    }
  }
  return acc;
}

void orc_constraints(int nt, const int32_t* t_i, const real* x, const real* Dm_inv,
                     const real* alpha, const real* lambda, real* C_out, real* gradC,
                     real* rhs, double* norm2) {
  ConsArgs a = {t_i, x, NULL, 0, Dm_inv, alpha, lambda, C_out, gradC, rhs, 0};
  *norm2 = chunked_sum(nt, cons_range, &a);
}

/* Two constraints per tet (arrays of 2 nt entries, 12 gradient floats each): constraint 2t
 * the ARAP constraint, 2t+1 the volume constraint (gradient never zeroed). */
void orc_constraints_vol(int nt, const int32_t* t_i, const real* x, const real* Dm_inv,
                         const real* alpha, const real* lambda, real* C_out, real* gradC,
                         real* rhs, double* norm2) {
  ConsArgs a = {t_i, x, NULL, 0, Dm_inv, alpha, lambda, C_out, gradC, rhs, 1};
  *norm2 = chunked_sum(nt, cons_range, &a);
}

static double dual_range(int lo, int hi, void* arg) {
  const ConsArgs* a = (const ConsArgs*)arg;
  double acc = 0.0;
  for (int t = lo; t < hi; ++t) {
    double p[4][3], g[4][3];
    tet_positions(a->t_i, t, a->x, a->dx, a->omega, p);
    const int64_t c0 = a->volume ? 2 * (int64_t)t : t;
    real c = (real)tet_C_grad(p, a->Dm_inv + 9 * (int64_t)t, g, 0);
    real rr = -(c + a->alpha[c0] * a->lambda[c0]);
    acc += (double)rr * (double)rr;
    if (a->volume) {
      real ch = (real)tet_vol_grad(p, a->Dm_inv + 9 * (int64_t)t, g, 0);
      real rh = -(ch + a->alpha[c0 + 1] * a->lambda[c0 + 1]);
      acc += (double)rh * (double)rh;
    }
  }
  return acc;
}

/* squared dual norm sum_t (C_t(x + omega dx) + alpha_t lambda_t)^2 */
double orc_dual_norm2(int nt, const int32_t* t_i, const real* x, const real* dx, real omega,
                      const real* Dm_inv, const real* alpha, const real* lambda) {
  ConsArgs a = {t_i, x, dx, omega, Dm_inv, alpha, lambda, NULL, NULL, NULL, 0};
  return chunked_sum(nt, dual_range, &a);
}

double orc_dual_norm2_vol(int nt, const int32_t* t_i, const real* x, const real* dx, real omega,
                          const real* Dm_inv, const real* alpha, const real* lambda) {
  ConsArgs a = {t_i, x, dx, omega, Dm_inv, alpha, lambda, NULL, NULL, NULL, 1};
  return chunked_sum(nt, dual_range, &a);
}

/* ---------------------------------------------------------------- A fill (CSR) */
static real dot3r(const real* a, const real* b) { return FMA(a[2], b[2], FMA(a[1], b[1], a[0] * b[0])); }

void orc_fill_A(int nt, const int32_t* indptr, const int32_t* indices, const int32_t* t_i,
                const real* invm, const real* gradC, const real* alpha, real* data) {
#pragma omp parallel for schedule(dynamic, 256) if (nt > ORC_PAR_MIN)
  for (int i = 0; i < nt; ++i) {
    const int32_t* a = t_i + 4 * (int64_t)i;
    const real* gi = gradC + 12 * (int64_t)i;
    real w[4] = {invm[a[0]], invm[a[1]], invm[a[2]], invm[a[3]]};
    for (int32_t n = indptr[i]; n < indptr[i + 1]; ++n) {
      int j = indices[n];
      real v;
      if (j == i) {
        v = w[0] * dot3r(gi, gi);
        v = FMA(w[1], dot3r(gi + 3, gi + 3), v);
        v = FMA(w[2], dot3r(gi + 6, gi + 6), v);
        v = FMA(w[3], dot3r(gi + 9, gi + 9), v);
        v = v + alpha[i];
      } else {
        const int32_t* b = t_i + 4 * (int64_t)j;
        const real* gj = gradC + 12 * (int64_t)j;
        v = 0;
        /* _intersect_tets order: slot of i outer, slot of j inner (cons/mgpbd.py:77-83) */
        for (int p = 0; p < 4; ++p)
          for (int q = 0; q < 4; ++q)
            if (a[p] == b[q]) v = FMA(w[p], dot3r(gi + 3 * p, gj + 3 * q), v);
      }
      data[n] = v;
    }
  }
}

void orc_cache_diag_inv(int n, const int32_t* indptr, const int32_t* indices, const real* data,
                        real* dinv) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) {
    real d = 0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
      if (indices[k] == i) { d = data[k]; break; }
    dinv[i] = RABS(d) > (real)1e-30 ? (real)1 / d : (real)0;
  }
}

void orc_spmv(int n, const int32_t* indptr, const int32_t* indices, const real* data,
              const real* x, real* y) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) {
    real s = 0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) s = FMA(data[k], x[indices[k]], s);
    y[i] = s;
  }
}

void orc_jacobi(int n, const int32_t* indptr, const int32_t* indices, const real* data,
                const real* dinv, const real* b, const real* x, real* x_new, real omega) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) {
    real s = 0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) s = FMA(data[k], x[indices[k]], s);
    x_new[i] = FMA(omega * dinv[i], b[i] - s, x[i]);
  }
}

/* coarse[c] = sum_k p_ki * fine[idx[k]] * p_lj (unit p: a plain sequential sum) */
void orc_refill(int nc, const int32_t* off, const int32_t* fine_idx, const real* pki,
                const real* plj, const real* fine, real* coarse) {
#pragma omp parallel for schedule(static) if (nc > ORC_PAR_MIN)
  for (int c = 0; c < nc; ++c) {
    real v = 0;
    for (int32_t k = off[c]; k < off[c + 1]; ++k) v = v + (pki[k] * fine[fine_idx[k]]) * plj[k];
    coarse[c] = v;
  }
}

/* ---------------------------------------------------------------- V-cycle + PCG */
typedef struct {
  int n;
  const int32_t *ip, *ix;   /* A_l CSR */
  const real *a, *dinv;
  const int32_t *pip, *pix; /* P_l  (n_l x n_{l+1}) */
  const real* pa;
  const int32_t *tip, *tix; /* P_l^T */
  const real* ta;
  real *b, *x, *xn, *r, *tmp;
  const double* dense;      /* coarsest only: dense column-major inverse, or NULL (sweeps) */
  real omega;               /* this level's Jacobi weight; <= 0: the cycle's common weight */
  const real* cheb;         /* Chebyshev (c1, c2) of smoothing steps 2..17, or NULL (Jacobi) */
} OrcLevel;

typedef struct {
  int nl, n_smooth;
  real omega;
  OrcLevel* L;
} OrcMG;

/* x_new = x + c1 (x - xp) + (c2 dinv) (b - A x): step >= 2 of the Chebyshev smoother */
static void orc_cheb_step(int n, const int32_t* indptr, const int32_t* indices, const real* data,
                          const real* dinv, const real* b, const real* x, const real* xp,
                          real* x_new, real c1, real c2) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) {
    real s = 0;
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) s = FMA(data[k], x[indices[k]], s);
    const real u = FMA(c1, x[i] - xp[i], x[i]);
    x_new[i] = FMA(c2 * dinv[i], b[i] - s, u);
  }
}

/* one smoothing phase of `count` steps from A->x (A->tmp keeps the previous iterate) */
static void smooth_phase(const OrcLevel* A, int count, real w) {
  const int n = A->n;
  real* prev = A->tmp;
  memcpy(prev, A->x, sizeof(real) * n);
  for (int s = 0; s < count; ++s) {
    const int step = 1 + s;
    if (A->cheb && step >= 2) {
      const int k = step - 2 < 15 ? step - 2 : 15;
      orc_cheb_step(n, A->ip, A->ix, A->a, A->dinv, A->b, A->x, prev, A->xn, A->cheb[2 * k],
                    A->cheb[2 * k + 1]);
    } else {
      orc_jacobi(n, A->ip, A->ix, A->a, A->dinv, A->b, A->x, A->xn, w);
    }
    memcpy(prev, A->x, sizeof(real) * n);
    memcpy(A->x, A->xn, sizeof(real) * n);
  }
}

static void vc(const OrcMG* mg, int l) {
  const OrcLevel* A = &mg->L[l];
  int n = A->n;
  const real w = A->omega > 0 ? A->omega : mg->omega;
  if (l == mg->nl - 1 && A->dense) {
    /* x = A^-1 b with FP64 accumulation */
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
    for (int i = 0; i < n; ++i) {
      double acc = 0;
      for (int j = 0; j < n; ++j) acc += A->dense[(int64_t)j * n + i] * (double)A->b[j];
      A->x[i] = (real)acc;
    }
    return;
  }
  if (l == mg->nl - 1) {
    int it = n < 50 ? n : 50;
    for (int s = 0; s < it; ++s) {
      orc_jacobi(n, A->ip, A->ix, A->a, A->dinv, A->b, A->x, A->xn, w);
      memcpy(A->x, A->xn, sizeof(real) * n);
    }
    return;
  }
  const OrcLevel* C = &mg->L[l + 1];
  smooth_phase(A, mg->n_smooth, w);
  orc_spmv(n, A->ip, A->ix, A->a, A->x, A->r);
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) A->r[i] = A->b[i] - A->r[i];
  orc_spmv(C->n, A->tip, A->tix, A->ta, A->r, C->b);
  memset(C->x, 0, sizeof(real) * C->n);
  vc(mg, l + 1);
  orc_spmv(n, A->pip, A->pix, A->pa, C->x, A->tmp);
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
  for (int i = 0; i < n; ++i) A->x[i] = FMA((real)1, A->tmp[i], A->x[i]);
  smooth_phase(A, mg->n_smooth, w);
}

typedef struct { const real *a, *b; } DotArgs;
static double dot_range(int lo, int hi, void* arg) {
  const DotArgs* d = (const DotArgs*)arg;
  double s = 0;
  for (int i = lo; i < hi; ++i) s += (double)d->a[i] * (double)d->b[i];
  return s;
}
static double dotd(int n, const real* a, const real* b) {
  DotArgs d = {a, b};
  return chunked_sum(n, dot_range, &d);
}

/* z = V-cycle(r) with x_0 = 0 */
void orc_vcycle(int nl, OrcLevel* L, int n_smooth, real omega, const real* r, real* z) {
  OrcMG mg = {nl, n_smooth, omega, L};
  memcpy(L[0].b, r, sizeof(real) * L[0].n);
  memset(L[0].x, 0, sizeof(real) * L[0].n);
  vc(&mg, 0);
  memcpy(z, L[0].x, sizeof(real) * L[0].n);
}

/* Returns cg_iters; info = {rr, bb, reason}. work: 4n reals (r, z, p, Ap). */
int orc_pcg(int nl, OrcLevel* L, int n_smooth, real omega, int maxiter, double tol,
            int relative, const real* b, real* x, real* work, double* info) {
  int n = L[0].n;
  real *r = work, *z = work + n, *p = work + 2 * (int64_t)n, *Ap = work + 3 * (int64_t)n;
  memset(x, 0, sizeof(real) * n);
  memcpy(r, b, sizeof(real) * n);
  double bb = dotd(n, b, b);
  double tol2 = relative ? tol * tol * bb : tol * tol;
  orc_vcycle(nl, L, n_smooth, omega, r, z);
  memcpy(p, z, sizeof(real) * n);
  double rz = dotd(n, r, z);
  int cg_iters = 0, reason = 0;
  double rr = 0;
  for (int k = 0; k < maxiter; ++k) {
    orc_spmv(n, L[0].ip, L[0].ix, L[0].a, p, Ap);
    double pAp = dotd(n, p, Ap);
    if (fabs(pAp) < 1e-30) { reason = 2; break; }
    real al = (real)(rz / pAp);
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
    for (int i = 0; i < n; ++i) {
      x[i] = FMA(al, p[i], x[i]);
      r[i] = FMA(-al, Ap[i], r[i]);
    }
    rr = dotd(n, r, r);
    if (rr < tol2) { cg_iters = k + 1; reason = 1; break; }
    orc_vcycle(nl, L, n_smooth, omega, r, z);
    double rz_new = dotd(n, r, z);
    if (fabs(rz) < 1e-30) { cg_iters = k + 1; reason = 3; break; }
    real be = (real)(rz_new / rz);
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
    for (int i = 0; i < n; ++i) p[i] = FMA(be, p[i], z[i]);
    rz = rz_new;
    cg_iters = k + 1;
  }
  if (info) { info[0] = rr; info[1] = bb; info[2] = reason; }
  return cg_iters;
}

/* ---------------------------------------------------------------- scatter */
/* lambda += dlam; dpos_v = sum_(t,slot) invm_v dlam_t g_t,slot accumulated in tet order */
void orc_scatter(int nt, int nv, const int32_t* t_i, const real* invm, const real* gradC,
                 const real* dlam, real* lambda, real* dpos) {
  memset(dpos, 0, sizeof(real) * 3 * (int64_t)nv);
  for (int t = 0; t < nt; ++t) {
    lambda[t] = lambda[t] + dlam[t];
    for (int l = 0; l < 4; ++l) {
      int v = t_i[4 * (int64_t)t + l];
      real s = invm[v] * dlam[t];
      for (int r = 0; r < 3; ++r)
        dpos[3 * (int64_t)v + r] = FMA(s, gradC[12 * (int64_t)t + 3 * l + r], dpos[3 * (int64_t)v + r]);
    }
  }
}

/* ---------------------------------------------------------------- CSR pattern */
/* cons/mgpbd_gpu.py:153-194 for large meshes: row t = every other tet sharing a vertex with
 * t, ascending, then t. Pass indices == NULL to get row lengths (into indptr[1..]) only. */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int64_t orc_pattern(int nt, int nv, const int32_t* t_i, int32_t* indptr, int32_t* indices) {
  int32_t* vptr = calloc((size_t)nv + 1, sizeof(int32_t));
  int32_t* vt = malloc(sizeof(int32_t) * 4 * (size_t)nt);
  int32_t* fill = malloc(sizeof(int32_t) * ((size_t)nv + 1));
  for (int64_t k = 0; k < 4 * (int64_t)nt; ++k) vptr[t_i[k] + 1]++;
  for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  memcpy(fill, vptr, sizeof(int32_t) * nv);
  for (int t = 0; t < nt; ++t)
    for (int l = 0; l < 4; ++l) vt[fill[t_i[4 * (int64_t)t + l]]++] = t;
  int cap = 1024;
  int32_t* buf = malloc(sizeof(int32_t) * cap);
  int64_t nnz = 0;
  if (!indices) indptr[0] = 0;
  for (int t = 0; t < nt; ++t) {
    int m = 0;
    for (int l = 0; l < 4; ++l) {
      int v = t_i[4 * (int64_t)t + l];
      int len = vptr[v + 1] - vptr[v];
      if (m + len > cap) { cap = 2 * (m + len); buf = realloc(buf, sizeof(int32_t) * cap); }
      memcpy(buf + m, vt + vptr[v], sizeof(int32_t) * len);
      m += len;
    }
    qsort(buf, m, sizeof(int32_t), cmp_i32);
    int u = 0;
    for (int k = 0; k < m; ++k)
      if ((k == 0 || buf[k] != buf[k - 1]) && buf[k] != t) buf[u++] = buf[k];
    if (indices) {
      memcpy(indices + indptr[t], buf, sizeof(int32_t) * u);
      indices[indptr[t] + u] = t;
    } else {
      indptr[t + 1] = indptr[t] + u + 1;
    }
    nnz += u + 1;
  }
  free(buf); free(fill); free(vt); free(vptr);
  return nnz;
}
