/*
 * mgpbd.h -- C ABI of libmgpbd.so, the MI355X (gfx950) backend of the MGPBD global XPBD solve.
 *
 * The reference implements this path in Python + Taichi (cons/mgpbd_gpu.py, cons/mgpcg_gpu.py,
 * cons/sparse_gpu.py, cons/tet_aggregation.py); its "FFI" is Taichi's kernel launch from
 * Python. This library replaces those kernels and host loops; the Python package
 * controllable_pbd_3d_amd binds it with ctypes behind the reference's class interface.
 * Each declaration cites the reference function it replaces.
 *
 * Conventions
 *  - Every function returns int status: 0 = ok, < 0 = error (MGPBD_E*); the message is in
 *    mgpbd_last_error() (thread-local). No exception crosses the ABI.
 *  - Pointers named *_dev are device pointers (hipMalloc / torch cuda tensors); all other
 *    pointers are host pointers. The library never frees caller memory.
 *  - Device work is queued on the context's stream (mgpbd_ctx_create / _set_stream) and is
 *    asynchronous unless stated; functions that return host values synchronise that stream.
 *  - Matrices on the device use the sliced layout "SELL-64x4": rows are binned 64 per slice
 *    (one gfx950 wavefront); slice s holds width_s = (slice_ptr[s+1]-slice_ptr[s])/64
 *    entries per row (a multiple of 4); entry k of row 64s+l is at
 *    slice_ptr[s] + 256*(k/4) + 4*l + (k%4), so a lane loads 4 entries (16 bytes) at a time.
 *    Rows shorter than width_s are padded with value 0 and column = the row itself.
 *    Within a row, entries keep the CSR order of the reference (level 0: ascending
 *    neighbours then the diagonal; coarse levels: ascending columns).
 */
#ifndef MGPBD_H
#define MGPBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPBD_ABI_VERSION 1

#define MGPBD_OK 0
#define MGPBD_EINVAL (-1)     /* invalid argument / shape */
#define MGPBD_EHIP (-2)       /* HIP runtime error */
#define MGPBD_ENONFINITE (-3) /* NaN/Inf met where a finite value is required */
#define MGPBD_EINTERNAL (-4)
#define MGPBD_ENOMEM (-5)

const char* mgpbd_last_error(void);
int mgpbd_abi_version(void);

/* ===================================================================================== */
/* Host-side setup (CPU, run once at solver init)                                         */
/* ===================================================================================== */

typedef struct mgpbd_pattern mgpbd_pattern;
/* CSR pattern of A = J M^-1 J^T + alpha: row t = vertex-sharing tets ascending, then t.
 * Replaces MGPBDSolverGPU._build_sparsity_pattern (cons/mgpbd_gpu.py:153-194). */
mgpbd_pattern* mgpbd_pattern_build(int n_tets, int n_verts, const int32_t* t_i);
int64_t mgpbd_pattern_nnz(const mgpbd_pattern* p);
int mgpbd_pattern_copy(const mgpbd_pattern* p, int32_t* indptr /*n_tets+1*/, int32_t* indices /*nnz*/);
void mgpbd_pattern_free(mgpbd_pattern* p);

/* Face adjacency with the neighbour order CPython's sets give the reference.
 * Replaces build_tet_face_adjacency (cons/tet_aggregation.py:11-37).
 * adj: 4*n_tets (-1 padded), deg: n_tets. */
int mgpbd_face_adjacency(int n_tets, const int32_t* t_i, int32_t* adj, int32_t* deg);

/* Greedy face-adjacent aggregation. Replaces tet_geometry_aggregation
 * (cons/tet_aggregation.py:40-127). */
int mgpbd_tet_aggregate(int n_tets, const int32_t* adj, const int32_t* deg, int max_agg,
                        int32_t* agg_ids, int32_t* n_agg);

/* Same growth rule on a general sorted CSR graph (deeper levels of the extended hierarchy;
 * no reference counterpart: the reference's coarse "virtual tets" never aggregate,
 * cons/tet_aggregation.py:199-208). */
int mgpbd_graph_aggregate(int n, const int32_t* ptr, const int32_t* idx, int max_agg,
                          int32_t* agg_ids, int32_t* n_agg);
/* pyamg's standard aggregation (the aggregation='pyamg' branch, cons/mgpbd_gpu.py:124-134;
 * pyamg itself is not vendored): pass 1 in node order makes a node whose neighbours are all
 * free an aggregate with them (a node without off-diagonal neighbours stays out, id -1);
 * pass 2 puts every free node into the pass-1 aggregate of its first such neighbour (adjacency
 * order); pass 3 makes each still-free node an aggregate with its free neighbours. ptr/idx:
 * the strength graph (pyamg's symmetric strength with theta = 0 keeps every stored entry of
 * A, so the matrix pattern itself). */
int mgpbd_standard_aggregate(int n, const int32_t* ptr, const int32_t* idx, int32_t* agg_ids,
                             int32_t* n_agg_out);
/* Extension (no reference counterpart): in node order, a node alone in its aggregate joins
 * the smallest aggregate among its neighbours' (ties: lowest id); aggregates are renumbered in
 * ascending order of their old ids with empty ones dropped; *n_agg_inout is updated. Applied
 * to the graph levels of the extended hierarchy, where greedy growth alone leaves most nodes
 * of the coarse graphs of irregular meshes as singletons. */
int mgpbd_graph_merge_singletons(int n, const int32_t* ptr, const int32_t* idx,
                                 int32_t* agg_ids, int32_t* n_agg_inout);
typedef struct mgpbd_graph mgpbd_graph;
mgpbd_graph* mgpbd_graph_coarsen(int n_f, const int32_t* ptr, const int32_t* idx,
                                 const int32_t* agg, int n_c);
int64_t mgpbd_graph_nnz(const mgpbd_graph* g);
int mgpbd_graph_copy(const mgpbd_graph* g, int32_t* ptr, int32_t* idx);
void mgpbd_graph_free(mgpbd_graph* g);

/* Galerkin coarse level for P[i, agg[i]] = 1 and its refill map.
 * Replaces GPUMultigridSolver.setup's P^T A P and contribution maps
 * (cons/mgpcg_gpu.py:46-99). values != NULL reproduces scipy's dropping of entries that
 * sum to 0.0 (value-dependent pattern); values == NULL gives the structural pattern. */
typedef struct mgpbd_galerkin mgpbd_galerkin;
mgpbd_galerkin* mgpbd_galerkin_build(int n_f, const int32_t* indptr, const int32_t* indices,
                                     const float* values, const int32_t* agg, int n_c);
int mgpbd_galerkin_sizes(const mgpbd_galerkin* g, int64_t* c_nnz, int64_t* n_contrib);
int mgpbd_galerkin_copy(const mgpbd_galerkin* g, int32_t* c_indptr, int32_t* c_indices,
                        int32_t* c_diag, int32_t* contrib_idx, int32_t* contrib_off);
void mgpbd_galerkin_free(mgpbd_galerkin* g);

/* vertex -> (4*tet + slot) incidence, ascending; used for the deterministic scatter that
 * replaces the atomics of _scatter_dlam_gpu (cons/mgpbd_gpu.py:330-343). */
int mgpbd_vertex_incidence(int n_tets, int n_verts, const int32_t* t_i, int32_t* vptr,
                           int32_t* vinc);

/* SELL-64x4 layout of a CSR matrix: slice_ptr (n_slices+1) and, per CSR position, its
 * position in the sliced arrays (sell_pos, nnz). No reference counterpart (layout only). */
int mgpbd_sell_layout(int n_rows, const int32_t* indptr, int32_t* slice_ptr);
int mgpbd_sell_map(int n_rows, const int32_t* indptr, const int32_t* slice_ptr,
                   int32_t* sell_pos);
/* Column array of the sliced layout (padding points at the row itself; rows past n_rows at
 * column 0), sell_idx sized slice_ptr[n_slices]. */
int mgpbd_sell_columns(int n_rows, const int32_t* indptr, const int32_t* indices,
                       const int32_t* slice_ptr, int32_t* sell_idx);
/* LDS windows for the windowed reduced-precision sweeps (no reference counterpart: the
 * reference's SpMV gathers x from global memory). Blocks of MGPBD_WINDOW_ROWS rows (the sweep
 * kernels' workgroup, ceil(64 n_slices / MGPBD_WINDOW_ROWS) blocks); for each, the sorted ids
 * of the MGPBD_WINDOW_SEG-column segments holding its rows' columns and own indices, in
 * segs[seg_ptr[b] .. seg_ptr[b+1]). seg_ptr has n_blocks + 1 entries; with segs NULL only
 * seg_ptr is filled (seg_ptr[n_blocks] = the size segs needs), else cap >= that size. */
#define MGPBD_WINDOW_ROWS 256
#define MGPBD_WINDOW_SEG 32
#define MGPBD_WINDOW_MAX_SEGS 384
int mgpbd_sell_window(int n_rows, const int32_t* indptr, const int32_t* indices,
                      int32_t* seg_ptr, int32_t* segs, int64_t cap);

/* ===================================================================================== */
/* Device context                                                                        */
/* ===================================================================================== */

typedef struct mgpbd_ctx mgpbd_ctx;
/* Binds to the current HIP device; `stream` is a hipStream_t (NULL = null stream). */
mgpbd_ctx* mgpbd_ctx_create(void* stream);
int mgpbd_ctx_set_stream(mgpbd_ctx* ctx, void* stream);
void mgpbd_ctx_destroy(mgpbd_ctx* ctx);
int mgpbd_ctx_sync(mgpbd_ctx* ctx);
/* Device memory the context itself holds (reduction partials, PCG state, smoother weights,
 * coarsest-level dense buffers), in bytes; caller-owned arrays are not counted. */
int mgpbd_ctx_device_bytes(mgpbd_ctx* ctx, int64_t* bytes);

/* Optional timing of the main kernels with HIP events on the context stream (no reference
 * counterpart; the reference times whole frames in utils/bench.py:24-55). enable resets the
 * counters; read synchronises and returns total milliseconds and launch counts per kind. */
#define MGPBD_PROF_SPMV_L0 0      /* PCG A p (+ p.Ap) on level 0 */
#define MGPBD_PROF_JACOBI_L0 1    /* weighted-Jacobi sweep on level 0 */
#define MGPBD_PROF_RESID_L0 2     /* V-cycle residual b - A x on level 0 */
#define MGPBD_PROF_FILL_A 3
#define MGPBD_PROF_CONSTRAINTS 4
#define MGPBD_PROF_REFILL 5
#define MGPBD_PROF_COARSE 6       /* sweeps and residuals on levels >= 1 */
#define MGPBD_PROF_KINDS 8
int mgpbd_ctx_profile(mgpbd_ctx* ctx, int enable);
int mgpbd_ctx_profile_read(mgpbd_ctx* ctx, double* ms_out, int64_t* count_out);

/* ===================================================================================== */
/* Per-tet / per-vertex kernels (cons/mgpbd_gpu.py)                                      */
/* ===================================================================================== */

/* Dm^-1 (row-major 3x3 per tet), |det Dm|/6, alpha = 1/(mu dt^2 vol), evaluated in FP64 and
 * rounded to FP32; mu_dev (per tet) overrides the scalar mu when not NULL.
 * Replaces _precompute_rest_shape (:136-151). */
int mgpbd_rest_shape(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev,
                     const float* v_ref_dev, const float* mu_dev, double mu, double dt,
                     float* Dm_inv_dev, float* rest_vol_dev, float* alpha_dev);

/* Volume constraint (extension; the reference's MGPBD solve has one ARAP constraint per
 * tet): tet t owns constraints 2t (ARAP) and 2t+1 (volume C_H = det F - 1, the hydrostatic
 * constraint of cons/deform3d.py:81-91). As mgpbd_rest_shape, with alpha (2 n_tets entries)
 * alpha[2t] = 1/(mu dt^2 vol) and alpha[2t+1] = 1/(lambda dt^2 vol), lambda the Lame
 * parameter (cons/vbd.py:44-47; lam_dev per tet overrides the scalar lam). */
int mgpbd_rest_shape_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev,
                         const float* v_ref_dev, const float* mu_dev, double mu,
                         const float* lam_dev, double lam, double dt, float* Dm_inv_dev,
                         float* rest_vol_dev, float* alpha_dev);

/* v += dt g (+ dt invm f_ext); v *= damp; x_old = x; x += dt v.
 * Replaces make_prediction (:206-213), which moves every vertex, pinned ones included;
 * hold_pinned != 0 (extension) keeps vertices with invm == 0 in place with zero velocity.
 * f_ext_dev (3 per vertex) may be NULL. */
int mgpbd_predict(mgpbd_ctx* ctx, int n_verts, float* v_p_dev, float* v_v_dev,
                  float* v_p_old_dev, const float* v_invm_dev, const float* f_ext_dev,
                  float gx, float gy, float gz, float dt, float damp, int hold_pinned);

/* v = damp (x - x_old) / dt. Replaces update_vel (:215-219). */
int mgpbd_update_vel(mgpbd_ctx* ctx, int n_verts, const float* v_p_dev,
                     const float* v_p_old_dev, float* v_v_dev, float dt, float damp);

/* C = ||sigma(F) - 1||, gradC (4 x vec3 per tet, 0 when C <= 1e-6), rhs = -(C + alpha lambda)
 * and ||rhs||^2 (FP64, written to *norm2_dev). F, its SVD, C and gradC are evaluated in FP64
 * from the FP32 positions and rounded once to FP32. Replaces _compute_C_and_gradC (:225-247),
 * _compute_rhs_gpu (:253-258) and _residual_norm_gpu (:260-267). */
int mgpbd_constraints(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev, const float* v_p_dev,
                      const float* Dm_inv_dev, const float* alpha_dev, const float* lambda_dev,
                      float* C_dev, float* gradC_dev, float* rhs_dev, double* norm2_dev);
/* Two constraints per tet (extension, see mgpbd_rest_shape_vol): C, gradC, rhs, alpha and
 * lambda hold 2 n_tets entries; constraint 2t as mgpbd_constraints, constraint 2t+1 the
 * volume constraint det F - 1 with gradient from cof(F) (never zeroed); ||rhs||^2 over
 * both. */
int mgpbd_constraints_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev,
                          const float* v_p_dev, const float* Dm_inv_dev, const float* alpha_dev,
                          const float* lambda_dev, float* C_dev, float* gradC_dev,
                          float* rhs_dev, double* norm2_dev);
/* The same constraint pass without a context and without allocating: no counterpart in the
 * reference. For callers that cannot call hipMalloc (HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). Same kernels and results as mgpbd_constraints, queued
 * on `stream` (a hipStream_t, NULL = the null stream). The ||rhs||^2 block partials live in
 * a file-scope device array of 2048 doubles (16 KiB), held by every GPU the library runs on
 * until the process ends; its address is looked up once per device on the first call there.
 * n_tets must be in [0, 2^28] (MGPBD_EINVAL otherwise). Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int mgpbd_constraints_noalloc(void* stream, int n_tets, const int32_t* t_i_dev,
                              const float* v_p_dev, const float* Dm_inv_dev,
                              const float* alpha_dev, const float* lambda_dev, float* C_dev,
                              float* gradC_dev, float* rhs_dev, double* norm2_dev);

/* A values on the level-0 sliced pattern plus 1/diag (0 when |diag| <= 1e-30).
 * Replaces _fill_A_kernel (:273-309) with the shared-vertex sum order of _intersect_tets
 * (cons/mgpbd.py:71-84) and cache_diag_inv (cons/sparse_gpu.py:107-121). */
int mgpbd_fill_A(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr_dev,
                 const int32_t* sell_idx_dev, const int32_t* row_len_dev,
                 const int32_t* t_i_dev, const float* v_invm_dev, const float* gradC_dev,
                 const float* alpha_dev, float* sell_val_dev, float* diag_inv_dev);
/* The same A fill (same values bit for bit) from per-slot shared-vertex descriptors
 * (extension): mgpbd_fill_A_desc writes, once per pattern, one 16-bit word per SELL slot
 * (desc_dev: padded slots, zero-initialised): the diagonal flag 0x8000, or the count of
 * vertices rows i and j share (bits 12-13) and, per shared vertex in ascending local slot a of
 * tet i, (a | b << 2) with b its slot in tet j; rows that share all four vertices (the two
 * constraints of one tet, mgpbd_constraints_vol) get 0x4000 | (slot in j of i's vertex a) << 2a;
 * mgpbd_fill_A_shared then gathers only row j's gradients at the shared vertices (12 bytes
 * each instead of its vertex ids and whole gradient, 64 bytes). */
int mgpbd_fill_A_desc(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr_dev,
                      const int32_t* sell_idx_dev, const int32_t* row_len_dev,
                      const int32_t* t_i_dev, uint16_t* desc_dev);
int mgpbd_fill_A_shared(mgpbd_ctx* ctx, int n_tets, const int32_t* slice_ptr_dev,
                        const int32_t* sell_idx_dev, const int32_t* row_len_dev,
                        const uint16_t* desc_dev, const int32_t* t_i_dev, const float* v_invm_dev,
                        const float* gradC_dev, const float* alpha_tilde_dev, float* val_dev,
                        float* diag_inv_dev);

/* lambda += dlambda; dpos_v = sum over (t, slot) at v of invm_v dlambda_t gradC_t,slot in
 * ascending (t, slot) order. Replaces _zero_dpos + _scatter_dlam_gpu (:330-354). */
int mgpbd_scatter(mgpbd_ctx* ctx, int n_tets, int n_verts, const int32_t* vptr_dev,
                  const int32_t* vinc_dev, const float* v_invm_dev, const float* gradC_dev,
                  const float* dlam_dev, float* lambda_dev, float* dpos_dev);

/* Dual-residual norms for the line search (cons/mgpbd_gpu.py:360-391): out[0] =
 * ||C(x) + alpha lambda||^2 from the stored C, out[1] = same at x + omega dpos (C
 * recomputed). FP64, device out (2 doubles). with_current: 0 out[1] only, 1 both, 2 out[0]
 * only (no constraint evaluation). */
int mgpbd_dual_norms(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev, const float* v_p_dev,
                     const float* dpos_dev, float omega, const float* Dm_inv_dev,
                     const float* alpha_dev, const float* lambda_dev, const float* C_dev,
                     int with_current, double* out_dev);

/* mgpbd_dual_norms with two constraints per tet (see mgpbd_constraints_vol). */
int mgpbd_dual_norms_vol(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev,
                         const float* v_p_dev, const float* dpos_dev, float omega,
                         const float* Dm_inv_dev, const float* alpha_dev,
                         const float* lambda_dev, const float* C_dev, int with_current,
                         double* out_dev);

/* x += omega dpos where invm > 0. Replaces _apply_dpos (:345-349). */
int mgpbd_apply_dpos(mgpbd_ctx* ctx, int n_verts, float* v_p_dev, const float* dpos_dev,
                     const float* v_invm_dev, float omega);

/* ===================================================================================== */
/* Sparse / BLAS-1 kernels (cons/sparse_gpu.py, cons/mgpcg_gpu.py)                       */
/* ===================================================================================== */

/* y = A x. Replaces GPUSparseMatrix.spmv (cons/sparse_gpu.py:123-132). */
int mgpbd_spmv(mgpbd_ctx* ctx, int n_rows, const int32_t* slice_ptr_dev,
               const int32_t* sell_idx_dev, const float* sell_val_dev, const float* x_dev,
               float* y_dev);
/* x_new = x + omega D^-1 (b - A x). Replaces jacobi_smooth (cons/sparse_gpu.py:134-144). */
int mgpbd_jacobi(mgpbd_ctx* ctx, int n_rows, const int32_t* slice_ptr_dev,
                 const int32_t* sell_idx_dev, const float* sell_val_dev,
                 const float* diag_inv_dev, const float* b_dev, const float* x_dev,
                 float* x_new_dev, float omega);
/* diag_inv[i] = 1/val[diag_pos[i]] (0 if |.| <= 1e-30 or diag_pos[i] < 0).
 * Replaces cache_diag_inv (cons/sparse_gpu.py:107-121). */
int mgpbd_diag_inv(mgpbd_ctx* ctx, int n_rows, const int32_t* diag_pos_dev,
                   const float* sell_val_dev, float* diag_inv_dev);
/* Coarse values: val_c[cpos[c]] = sum_{k in [off[c], off[c+1])} val_f[fpos[k]] in order.
 * Replaces _fill_coarse_values_kernel (cons/mgpcg_gpu.py:183-198) for unit P entries. */
int mgpbd_coarse_refill(mgpbd_ctx* ctx, int n_coarse_nnz, const int32_t* off_dev,
                        const int32_t* fpos_dev, const int32_t* cpos_dev,
                        const float* fine_val_dev, float* coarse_val_dev);
/* b_c[I] = sum_{i in I, ascending} r[i] (P^T r for unit piecewise-constant P).
 * Replaces the P^T spmv of vcycle (cons/mgpcg_gpu.py:243). */
int mgpbd_restrict(mgpbd_ctx* ctx, int n_coarse, const int32_t* mem_ptr_dev,
                   const int32_t* mem_dev, const float* r_dev, float* b_c_dev);
/* x[i] += x_c[agg[i]]. Replaces P spmv + axpy of vcycle (cons/mgpcg_gpu.py:252-253). */
int mgpbd_prolong(mgpbd_ctx* ctx, int n_fine, const int32_t* agg_dev, const float* x_c_dev,
                  float* x_dev);

/* BLAS-1 (cons/sparse_gpu.py:17-69). Reductions are FP64 and deterministic and write their
 * result to a device double. */
int mgpbd_dot(mgpbd_ctx* ctx, int n, const float* a_dev, const float* b_dev, double* out_dev);
int mgpbd_axpy(mgpbd_ctx* ctx, int n, float alpha, const float* x_dev, float* y_dev);
int mgpbd_xpay(mgpbd_ctx* ctx, int n, float alpha, const float* x_dev, float* y_dev);
int mgpbd_copy(mgpbd_ctx* ctx, int n, const float* src_dev, float* dst_dev);
int mgpbd_fill(mgpbd_ctx* ctx, int n, float value, float* x_dev);
int mgpbd_scale(mgpbd_ctx* ctx, int n, float alpha, float* x_dev);
int mgpbd_negate_and_add(mgpbd_ctx* ctx, int n, const float* b_dev, float* r_dev);

/* ===================================================================================== */
/* Multigrid-preconditioned CG (cons/mgpcg_gpu.py GPUMultigridSolver)                    */
/* ===================================================================================== */

#define MGPBD_MAX_LEVELS 12
#define MGPBD_COARSE_JACOBI 0   /* min(coarse_sweeps, n) weighted-Jacobi sweeps from 0 */
#define MGPBD_COARSE_CHOLESKY 1 /* exact FP64 solve with the dense inverse (extension) */
#define MGPBD_MAX_DENSE 4096    /* largest coarsest level for MGPBD_COARSE_CHOLESKY */

/* (Re)declare the hierarchy: n_levels levels, level 0 the finest. */
int mgpbd_mg_configure(mgpbd_ctx* ctx, int n_levels, int n_smooth, float omega,
                       int coarse_mode, int coarse_sweeps);
/* Level l operator and transfer to level l+1 (agg/mem_ptr/mem NULL on the coarsest).
 * work_dev: 8 * 64 * ceil(n_rows / 64) floats owned by the caller (b, x, x', r of the level,
 * two W-cycle buffers and the spectral smoother's power-iteration vectors, each padded to
 * whole 64-row slices).
 * diagonal_only != 0 declares that every row holds only its diagonal entry (then repeated
 * Jacobi sweeps run as one elementwise pass with identical arithmetic). */
int mgpbd_mg_set_level(mgpbd_ctx* ctx, int level, int n_rows, const int32_t* slice_ptr_dev,
                       const int32_t* sell_idx_dev, const float* sell_val_dev,
                       const float* diag_inv_dev, const int32_t* agg_dev,
                       const int32_t* mem_ptr_dev, const int32_t* mem_dev, float* work_dev,
                       int diagonal_only);
/* Coarsest level with MGPBD_COARSE_CHOLESKY: assemble it densely in FP64 and keep its
 * inverse (library FP64 GEMM kernels, no vendor library): a full inversion runs Newton-Schulz
 * X <- X + X (I - A X) from X = I / ||A||_inf until ||I - A X||_F stops falling (below 1e-10,
 * or at its rounding floor below 0.1), checked by the host every four steps (it synchronises;
 * MGPBD_ENONFINITE when it does not converge within 64 steps: A not positive definite); a
 * refresh updates the previous inverse without a wait (mgpbd_mg_set_coarse_update). Call
 * after the coarsest values change. No reference counterpart (the reference smooths the
 * coarsest level, mgpcg_gpu.py:216-224). */
int mgpbd_mg_factor_coarsest(mgpbd_ctx* ctx);
/* Registers the coarse refill of level `level` + 1 from level `level` (the arguments of
 * mgpbd_coarse_refill) with the coarse level's diagonal positions and 1/diag buffer, for
 * mgpbd_mg_refresh. n_entries = 0 registers a level whose values come from elsewhere (the
 * factored level 1). Device pointers stay owned by the caller. */
int mgpbd_mg_set_refill(mgpbd_ctx* ctx, int level, int n_entries, const int32_t* off_dev,
                        const int32_t* fpos_dev, const int32_t* cpos_dev,
                        const float* fine_val_dev, float* coarse_val_dev, int n_coarse,
                        const int32_t* coarse_diag_pos_dev, float* coarse_dinv_dev);
/* One call for update_all_coarse_matrices (cons/mgpcg_gpu.py:183-212) after level 0's values
 * changed: per level the coarse refill (mgpbd_mg_factored_refill for a factored level 1)
 * and 1/diag, the reduced-precision copies of levels half_from .. n_levels-1 (a caller that
 * converted level 0 early passes 1) and, with refresh != 0, the coarsest inverse
 * (MGPBD_COARSE_CHOLESKY) and the spectral smoother's estimates; refresh = 0 keeps both
 * (the caller knows the matrices changed little; the kept inverse is SPD). coarse = 0 keeps
 * levels 1 .. n_levels-1 as they are (extension, no reference counterpart: the caller knows
 * level 0 changed little since they were built, e.g. within one substep's Newton loop; the
 * V-cycle stays an SPD preconditioner of level 0's current matrix, which the PCG solves):
 * then only level 0's reduced-precision data (half_from = 0) and, with refresh, the spectral
 * estimates are refreshed. Queues only (a Cholesky refactorization synchronises, as in
 * mgpbd_mg_factor_coarsest). */
int mgpbd_mg_refresh(mgpbd_ctx* ctx, int refresh, int half_from, int coarse);
/* Lagged hierarchy (extension; no reference counterpart: the reference rebuilds every level at
 * every Newton iteration, cons/mgpcg_gpu.py:183-212). Records whether levels 1.. are older
 * than level 0's current values; mgpbd_mg_refresh sets it itself (coarse = 0: lagged), a
 * caller that refills the levels through the separate entry points says so here. While it is
 * set, a PCG solve that is still running after max(8, 2 f + 2) iterations, f the iteration
 * count of the last solve on a current hierarchy, rebuilds levels 1.. (refill maps, 1/diag,
 * coarsest inverse, reduced-precision copies) and resumes from its iterate, once per solve:
 * a hierarchy gone stale within a substep costs a bounded number of iterations instead of
 * multiplying them. */
int mgpbd_mg_set_coarse_lagged(mgpbd_ctx* ctx, int lagged);
/* With newton_schulz != 0 (the default) a refresh of an already inverted coarsest level
 * takes two Newton-Schulz steps X <- X + X (I - A X) from the previous inverse and the device
 * keeps the result when ||I - A X||_F < 0.1 afterwards (then every eigenvalue of X A lies
 * within 0.1 of 1 and the V-cycle stays an SPD preconditioner). Otherwise, or when the
 * iteration diverged, the previous inverse stays for the solves that follow (SPD, so the
 * V-cycle stays SPD) and the next refresh inverts from scratch; the host learns the outcome at
 * that refresh, without waiting for it in this one. newton_schulz = 0: every refresh inverts
 * from scratch. */
int mgpbd_mg_set_coarse_update(mgpbd_ctx* ctx, int newton_schulz);
/* Counts since the context was created: Newton-Schulz refresh steps kept, full inversions. */
int mgpbd_mg_coarse_stats(mgpbd_ctx* ctx, int64_t* newton_schulz_steps, int64_t* refactors);
/* Smoothing steps per phase on one level (extension; the reference uses n_smooth on every
 * level, cons/mgpcg_gpu.py:232-246): steps in [1, 16], 0 restores n_smooth. Coarse levels are
 * cheap in bytes but each step is a launch, so fewer steps there can pay. */
int mgpbd_mg_set_level_smoothing(mgpbd_ctx* ctx, int level, int steps);
/* Cycle shape (extension; no reference counterpart: cons/mgpcg_gpu.py:214-259 is a V-cycle
 * adding the plain coarse correction). cycle 1 = V, 2 = W (each level below the finest that is
 * not the coarsest is cycled twice, the second time on its residual); overcorrect scales the
 * prolongated coarse correction (1 = the reference). The defaults are 1 and 1. */
int mgpbd_mg_set_cycle(mgpbd_ctx* ctx, int cycle, float overcorrect);
/* Launch fusion of the reduced-precision V-cycle (extension; no reference counterpart: the
 * reference prolongs in kernels of their own, cons/mgpcg_gpu.py:251-253). enable 1 (the
 * default): the dense coarsest product (coarse_mode Cholesky) also adds each coarse row's
 * correction to the members of its aggregate on the level above, and the PCG's update writes
 * the scaled level 0's rhs isd r and first sweep (one launch fewer per V-cycle and per
 * iteration); 0: separate launches. Both give the same iterates bit for bit. */
int mgpbd_mg_set_launch_fusion(mgpbd_ctx* ctx, int enable);
/* Smoother (extension). mode 0 = the reference's fixed weighted Jacobi (the omega of
 * mgpbd_mg_configure on every level, cons/mgpcg_gpu.py:232-246). mode 1 = spectral Jacobi:
 * level l relaxes with omega_l = target / (safety lambda_l), lambda_l the largest eigenvalue
 * of D^-1 A_l estimated by power iteration, warm-started (12 iterations for the first four
 * updates; then 8 after an update whose last two estimates differed by more than 2 %, else
 * power_iters; an unsettled estimate is raised by a quarter); omega_l lambda_max < 2 keeps the V-cycle an SPD
 * preconditioner, which the fixed 2/3 loses once lambda_max(D^-1 A) > 3. mode 2 = Chebyshev:
 * every smoothing phase of a level above the coarsest applies the Chebyshev polynomial of
 * D^-1 A_l on [safety lambda_l / cheb_ratio, safety lambda_l] (step 1 a Jacobi sweep with
 * 1 / theta, later steps x + c1 (x - x_prev) + c2 D^-1 (b - A x)); the same polynomial pre and
 * post keeps the cycle symmetric. Weights and coefficients live in device memory, so captured
 * PCG graphs need no recapture when they change. */
int mgpbd_mg_set_smoother(mgpbd_ctx* ctx, int mode, float target, float safety, int power_iters,
                          float cheb_ratio);
/* Re-estimates lambda_l and omega_l for the current level values (call after the coarse refill;
 * no-op in mode 0). lambda_out (host, n_levels doubles, may be NULL) receives the estimates
 * (then the call synchronises). */
int mgpbd_mg_update_smoother(mgpbd_ctx* ctx, double* lambda_out);
/* Fixed per-level weighted-Jacobi weights (extension): level l relaxes with omega[min(l, n-1)]
 * (host array, each in (0, 2)) instead of the common omega of cons/mgpcg_gpu.py:232-246; no
 * spectral estimates are made. For matrices whose lambda_max(D^-1 A) exceeds 3 on some level
 * (the volume constraint's rows), where the common 2/3 makes the V-cycle indefinite. Reset by
 * mgpbd_mg_set_smoother. */
int mgpbd_mg_set_level_weights(mgpbd_ctx* ctx, int n, const float* omega);
/* The weight each level relaxes with (host, n_levels floats) and, when cheb_out is given,
 * the Chebyshev (c1, c2) of steps 2..17 per level (n_levels x 32 floats, zero unless mode 2);
 * synchronises. */
int mgpbd_mg_get_smoother(mgpbd_ctx* ctx, float* omega_out, float* cheb_out);
/* Reduced-precision V-cycle (extension; no reference counterpart). Every level is used in its
 * scaled form D^-1/2 A D^-1/2 (unit diagonal), values stored FP16 or BF16 in the level's
 * SELL positions, products accumulated in FP32; the outer PCG SpMV stays FP32. flexible != 0
 * selects the Polak-Ribiere beta (flexible PCG) for the variable preconditioner. */
#define MGPBD_PREC_FP32 0
#define MGPBD_PREC_FP16 1
#define MGPBD_PREC_BF16 2
int mgpbd_mg_set_precision(mgpbd_ctx* ctx, int precision, int flexible);
/* Per-level buffers, exactly one of: hpacked_dev (uint32 per SELL slot, (col - row) << 16 |
 * half bits; only valid when |col - row| < 32768 on the whole level) or hval_dev (uint16 per
 * SELL slot, with the level's int32 columns); sd_dev / isd_dev (n_rows floats, written by
 * mgpbd_mg_convert_half), diag_pos_dev (SELL position of each row's diagonal). Level 0 may
 * pass neither buffer when it runs factored (mgpbd_mg_set_level0_factored). */
int mgpbd_mg_set_level_half(mgpbd_ctx* ctx, int level, uint16_t* hval_dev, uint32_t* hpacked_dev,
                            float* sd_dev, float* isd_dev, const int32_t* diag_pos_dev);
/* Switches a level to the windowed kernels (no reference counterpart; same results as the
 * gathering kernels, bit for bit): each workgroup stages its block's window of x in LDS
 * (seg_ptr_dev / segs_dev from mgpbd_sell_window on the level's CSR, n_blocks blocks, largest
 * window max_segs <= MGPBD_WINDOW_MAX_SEGS segments). On a packed level (mgpbd_mg_set_level_half
 * with hpacked) the words' high halves become LDS byte offsets (rewritten here;
 * mgpbd_mg_convert_half keeps them) for the reduced-precision sweeps and power steps; with
 * woff_dev (uint16 per SELL slot, filled here) the level-0 FP32 kernels of the PCG (A p, the
 * FP64 residual check) read x from the window too. At least one of the two must be present.
 * The arrays must stay valid while the level is used; mgpbd_mg_set_level / _set_level_half
 * reset the window. Synchronises the context's stream. */
int mgpbd_mg_set_level_window(mgpbd_ctx* ctx, int level, const int32_t* seg_ptr_dev,
                              const int32_t* segs_dev, int n_blocks, int max_segs,
                              uint16_t* woff_dev);
/* Factored level 0 for the reduced-precision solve (extension; no reference counterpart --
 * the reference assembles level 0 as the CSR of cons/mgpbd_gpu.py:273-324 every Newton
 * iteration and its PCG multiplies that CSR, cons/mgpcg_gpu.py:261-329). The level-0 matrix
 * is A = J W J^T + diag(alpha~) with J's rows the tets' vertex gradients, so
 *  - the V-cycle's level-0 sweeps, residuals and power steps apply its scaled form
 *    K W' K^T + diag(a^), K_tk = c gradC_tk / sqrt(A_tt), W' = diag(invm) / c^2,
 *    a^_t = alpha~_t / A_tt, c^2 = max invm (read from v_invm here, once), from K stored
 *    FP16/BF16 (two passes: per-vertex gather of K x times w'_v, per-tet product); K of a
 *    tet's fourth vertex is the rounded negative sum of the first three (a tet constraint's
 *    gradients sum to zero), so the tet order stores nine values per tet;
 *  - the PCG's A p, its guessed-start and restart residuals and the FP64 check apply A itself
 *    with FP64 products and sums of the FP32 gradients (the exact operator of those
 *    gradients, where the assembled FP32 entries round each A_ij once);
 *  - mgpbd_mg_convert_half takes A_tt = sum_k invm |gradC_tk|^2 + alpha~_t, the scales and
 *    K from the arrays' current contents and writes 1 / A_tt to dinv.
 * So level 0's stored values are not read while it runs factored (they need not be current),
 * and its half values need not exist (mgpbd_mg_set_level_half(ctx, 0, NULL, NULL, sd, isd,
 * diag_pos) sets the scale arrays only). Caller-owned device arrays, valid while used: t_i
 * (int32[4 n_tets], 16-byte aligned), v_invm (n_verts), gradC (float[12 n_tets], 16-byte
 * aligned), alpha_tilde (n_tets); the vertex-sliced incidence, four lanes per vertex: slot k of
 * vertex 16 s + l at vs_ptr[s] + 64 (k / 4) + 4 l + k % 4 for k < vs_cnt[v], holding vs_inc =
 * 4 tet + local vertex index; outputs vs_k (2 uint32 per slot), vs_g (4 floats per slot,
 * 16-byte aligned), tk (18 bytes per tet: a 16-byte plane of n_tets, then a 2-byte plane;
 * 16-byte aligned), ahat (n_tets floats), u (4 floats per vertex,
 * 16-byte aligned), u64 (4 doubles per vertex, 32-byte aligned), dinv (n_tets floats). Needs
 * a reduced precision (mgpbd_mg_set_precision). Synchronises the device (the read of
 * v_invm). t_i NULL switches the factored operator off.
 * Level 0 must be set (mgpbd_mg_set_level) and its scale arrays given
 * (mgpbd_mg_set_level_half) first; setting level 0 again switches it off. */
int mgpbd_mg_set_level0_factored(mgpbd_ctx* ctx, int n_verts, const int32_t* t_i,
                                 const float* v_invm, const float* gradC,
                                 const float* alpha_tilde, const int32_t* vs_ptr,
                                 const int32_t* vs_cnt, const int32_t* vs_inc, uint32_t* vs_k,
                                 float* vs_g, uint32_t* tk, float* ahat, float* u, double* u64,
                                 float* dinv);
/* 16-bit slot offsets for the factored level 0's vertex passes (no reference counterpart):
 * tbase_dev (n_verts int32, device) holds a base tet per vertex such that every tet at the
 * vertex lies in [tbase, tbase + 65535]; the per-iteration conversion then writes each slot's
 * tet - tbase into the unused fourth word of its K and FP32 gradient, and the V-cycle's and
 * the PCG's vertex passes read no incidence list (12 -> 8 bytes per slot). NULL switches it
 * off; it takes effect at the next mgpbd_mg_convert_half. rows_uniform (0..8, 0 = off)
 * declares that the vertex slices set with mgpbd_mg_set_level0_factored are uniform: every
 * slice rows_uniform rows of 64 slots (vs_ptr[s] = 64 rows_uniform s) with the unused slots
 * of vs_k / vs_g zero; the vertex passes then compute each lane's slot positions instead of
 * loading the slice pointer and count (same sums bit for bit: the unused slots add exact
 * zeros). With tbase_dev the context allocates (once, 8 bytes per vertex; counted by
 * mgpbd_ctx_device_bytes) the {base tet, w'} words that the uniform vertex pass reads, which
 * mgpbd_mg_convert_half fills. */
int mgpbd_mg_set_level0_offsets(mgpbd_ctx* ctx, const int32_t* tbase_dev, int rows_uniform);
/* Packed row vertices for the factored level 0's tet passes (no reference counterpart): one
 * 64-bit word per level-0 row (device, 8-byte aligned) holding the row's first vertex in the
 * low base_bits bits and the other three as signed diff_bits-bit differences from it, in
 * order above (base_bits + 3 diff_bits <= 64). The V-cycle's tet pass, the power steps and
 * the PCG's tet pass then read 8 bytes per row instead of t_i's 16 (same vertices, same
 * results bit for bit). NULL switches it off (t_i again). */
int mgpbd_mg_set_level0_packed_rows(mgpbd_ctx* ctx, const uint64_t* rows_dev, int base_bits,
                                    int diff_bits);
/* Level 1 from the gradients with the factored level 0 (extension; replaces the level 0 -> 1
 * refill of cons/mgpcg_gpu.py:183-212 when level 0 is not assembled): with P[i, agg[i]] = 1,
 * A1 = G W G^T + P^T diag(alpha~) P, G_(I,v) = sum of gradC over the tets of aggregate I
 * holding vertex v. Per (aggregate, vertex) pair p < n_pairs: its incidences (4 tet + local
 * vertex) gmem[gptr[p] .. gptr[p+1]) and vertex gvert[p]; G (4 floats per pair, 16-byte
 * aligned) is written by mgpbd_mg_factored_refill. Per level-1 entry c < n_entries: the
 * pairs (a, b) of the vertices its two aggregates share, terms[2 m], terms[2 m + 1] for m in
 * [tptr[c], tptr[c+1]) (8-byte aligned), over the entries (I, J) with I <= J only (A1 is
 * symmetric): its SELL position cpos[c] in c_val (level 1's values), its transpose's
 * cpos_t[c] (the same on the diagonal), and cagg[c] = the aggregate when c is a diagonal
 * entry, else -1 (the diagonal adds the aggregate's alpha~ over level 0's member lists).
 * gptr NULL switches it off. */
int mgpbd_mg_set_level1_factored(mgpbd_ctx* ctx, int n_pairs, const int32_t* gptr,
                                 const int32_t* gmem, const int32_t* gvert, float* G,
                                 int n_entries, const int32_t* tptr, const int32_t* terms,
                                 const int32_t* cpos, const int32_t* cpos_t,
                                 const int32_t* cagg, float* c_val);
/* Level 1's values from the current gradients and alpha~ (FP64 sums rounded once; padding
 * slots untouched). */
int mgpbd_mg_factored_refill(mgpbd_ctx* ctx);
/* lambda += dlambda and dpos_v = invm_v sum_t dlambda_t gradC_tv with the factored level 0's
 * vertex-order gradients (current after mgpbd_mg_convert_half): the scatter of
 * _scatter_dlam_gpu (cons/mgpbd_gpu.py:330-343) with four lanes per vertex (the summation
 * order differs from mgpbd_scatter's). Device arrays: dlam, lambda (n_tets), dpos (3 n_verts). */
int mgpbd_mg_factored_scatter(mgpbd_ctx* ctx, const float* dlam_dev, float* lambda_dev,
                              float* dpos_dev);
/* Recomputes scales and reduced-precision values of every level from the FP32 values; call
 * after the values change (after the coarse refill). */
int mgpbd_mg_convert_half(mgpbd_ctx* ctx);
/* The same for levels first .. last-1 only (extension; e.g. the factored level 0 first, so
 * that work needing its scales can start on another stream while the coarse levels refill). */
int mgpbd_mg_convert_half_levels(mgpbd_ctx* ctx, int first, int last);
/* r = b - A_l x with the operator the V-cycle applies on level l (no reference counterpart;
 * the reference computes this residual inside vcycle, cons/mgpcg_gpu.py:230-234): the FP32
 * level, or in reduced precision its scaled form (stored half values, or the factored level
 * 0) with b, x, r in the scaled variables. Device arrays of n_l floats; x != r. */
int mgpbd_mg_level_residual(mgpbd_ctx* ctx, int level, const float* b_dev, const float* x_dev,
                            float* r_dev);
/* x = M^-1 b with one V-cycle from x = 0. Replaces vcycle(0) (cons/mgpcg_gpu.py:214-259). */
int mgpbd_mg_vcycle(mgpbd_ctx* ctx, const float* b_dev, float* x_dev);
/* Measurement (no reference counterpart): level-0 weighted-Jacobi sweeps
 * y = x + omega D^-1 (b - A x) with the kernels the V-cycle launches, on the context stream.
 * The first kernel is the factored level's vertex pass (or the whole sweep of a stored level),
 * the second the factored tet pass (none for a stored level). flush_floats = 0: `reps`
 * (<= 1000) first kernels back to back, then `reps` second kernels, then `reps` sweeps, with
 * HIP events only between the three batches; ms_out[0..2] = per-launch time of each batch
 * (kernel plus dispatch gap). flush_floats > 0: before each of `reps` sweeps the flush buffer
 * is rewritten (scale by 1), evicting the caches, and events around each kernel give
 * ms_out[0] first kernel, ms_out[1] second, ms_out[2] sweep (averages). */
int mgpbd_mg_probe_level0(mgpbd_ctx* ctx, int reps, const float* b_dev, const float* x_dev,
                          float* y_dev, float* flush_dev, int64_t flush_floats, double* ms_out);
/* PCG for A x = b, x from 0 (the reference's GPUMultigridSolver.solve, cons/mgpcg_gpu.py:
 * 261-329: stop when ||r||^2 < tol^2, breakdown on |pAp| < 1e-30 and on the previous
 * |rz| < 1e-30). relative != 0 (extension) stops on ||r|| < tol ||b||. work_dev: 4*n floats
 * (r, z, p, Ap), 6*n floats (8-byte aligned) with the FP64 iterate. Synchronises; *iters_out = completed iterations (reference count);
 * info_out (host, 4 doubles) = {||r||^2, ||b||^2, stop reason, FP64 ||b - A x||^2 when
 * verification is on, else -1}. */
int mgpbd_pcg_solve(mgpbd_ctx* ctx, const float* b_dev, float* x_dev, float* work_dev,
                    int maxiter, double tol, int relative, int* iters_out, double* info_out);
/* PCG launch mode (no reference counterpart): with enable != 0 (default) the iterations run
 * as a captured HIP graph of `poll` iterations on a library-owned stream ordered after the
 * context stream; the host checks the done flag after every `poll` iterations (default 4).
 * Iterations past convergence are skipped on the device by the done flag. */
int mgpbd_ctx_use_graphs(mgpbd_ctx* ctx, int enable, int poll);
/* Converged-solution verification (extension; the reference stops on the recurrence residual,
 * cons/mgpcg_gpu.py:306-309): with enable != 0, every solve computes ||b - A x|| of the
 * returned x with FP64 products and sums; when the recurrence residual met the tolerance but
 * that misses it, the PCG restarts from x (at most 3 times, within maxiter) and the stop
 * reason is 0 when it never verifies. Iteration counts include the restarts. enable = 2 computes the FP64 residual of
 * every returned x (converged or not) without restarting. */
int mgpbd_pcg_set_verify(mgpbd_ctx* ctx, int enable);
/* FP64 iterate (extension; the reference's PCG vectors are FP32, cons/mgpcg_gpu.py:264-272):
 * with enable != 0 the PCG accumulates x in FP64 (x += alpha p with the FP64 alpha), the outer
 * A p takes FP64 products and sums (rounded once to FP32), and a verification restart starts
 * from r = b - A x computed the same way from the FP64 iterate; x_dev receives the FP32
 * rounding of the iterate. The rounding errors of an FP32 x (one ulp of x per update, each
 * worth up to kappa(A) eps in the relative residual) otherwise bound the attainable residual
 * on stiff, high-contrast systems. Needs the 6*n-float work buffer. Default off. */
int mgpbd_pcg_set_fp64_iterate(mgpbd_ctx* ctx, int enable);
/* Stall rescue (extension; spectral and Chebyshev smoothers only): with enable != 0, a PCG
 * round whose squared recurrence residual falls by less than 1 / stall_ratio over 8
 * iterations (stall_ratio 0.25: less than 2x in the norm; checked from the 16th iteration
 * on, at the host's polls) stops, every level's spectral estimate is recomputed
 * from a fresh start vector (12 power steps) with the safety factor times boost (in [1, 4]),
 * and the PCG restarts from its iterate, at most once per solve, within maxiter. Under
 * abrupt loading (a new spring drag) the warm-started estimates can lag the largest
 * eigenvalue; the polynomial smoother then amplifies the modes above its interval and the
 * V-cycle stops contracting. Default off. */
int mgpbd_pcg_set_rescue(mgpbd_ctx* ctx, int enable, float boost, double stall_ratio);
/* Cumulative counts over the context's solves: rescues (stall and lagged-hierarchy) and
 * verification restarts. */
int mgpbd_pcg_counters(mgpbd_ctx* ctx, int64_t* rescues, int64_t* restarts);
/* Start criterion of the warm-started solves (extension): 0 (the default) the candidate start
 * (guess, alternative, zero) with the smallest residual ||b - A x0||; 1 the one with the lowest
 * CG energy phi(x0) = x0.A x0 / 2 - b.x0, i.e. the smallest A-norm error, which is what bounds
 * CG's iterations (a start with a small residual can carry a large smooth error). */
int mgpbd_pcg_set_start_criterion(mgpbd_ctx* ctx, int energy);
/* Iteration limit of the lagged-hierarchy rescue (extension; see mgpbd_mg_set_coarse_lagged):
 * -1 (the default) max(8, 2 f + 2), 0 off, k > 0 after k iterations. */
int mgpbd_pcg_set_lag_rescue(mgpbd_ctx* ctx, int limit);
/* Cumulative count of the lagged-hierarchy rescues alone (extension; see
 * mgpbd_mg_set_coarse_lagged). */
int mgpbd_pcg_lag_rescues(mgpbd_ctx* ctx, int64_t* lag_rescues);

/* ---- Warm start (extension; no reference counterpart: the reference zeroes lambda every
 * substep, cons/mgpbd_gpu.py:402, and starts every PCG from dlambda = 0,
 * cons/mgpbd_gpu.py:426 and cons/mgpcg_gpu.py:271-272) ---- */
/* PCG for A x = b from a guess: x_dev holds the guess on entry, x_alt_dev (or NULL) a second
 * one (it must not alias x_dev). The start with the smallest ||b - A x0|| among the guess,
 * the alternative and zero is taken on the device (the guess wins a tie with the
 * alternative, zero wins ties and non-finite residuals); from x0 = 0 the arithmetic is that of
 * mgpbd_pcg_solve. The stopping test is that of mgpbd_pcg_solve (||r|| < tol, or
 * ||r|| < tol ||b|| with relative != 0), so every start reaches the same tolerance.
 * info_out (host, 8 doubles) = {||r||^2, ||b||^2, stop reason, ||r0||^2 of the start taken,
 * start (0 zero, 1 guess, 2 alternative), ||b - A x_guess||^2, ||b - A x_alt||^2, FP64
 * ||b - A x||^2 when verification is on (mgpbd_pcg_set_verify), else -1}. */
int mgpbd_pcg_solve_guess(mgpbd_ctx* ctx, const float* b_dev, float* x_dev,
                          const float* x_alt_dev, float* work_dev, int maxiter, double tol,
                          int relative, int* iters_out, double* info_out);

/* Spring drag loading (extension): f_ext = 0 except f_ext[vertex] = k (target - x[vertex]).
 * Used with mgpbd_predict's f_ext. */
int mgpbd_spring_drag(mgpbd_ctx* ctx, int n_verts, const float* v_p_dev, float* f_ext_dev,
                      int vertex, float tx, float ty, float tz, float k);

/* ---- Spatial adaptivity (the callers of the solve in cons/adaptive_zone.py and
 * cons/mgpbd_adaptive.py). The zone is one int32 flag per tet (1 active, 0 passive) on the
 * device; all entry points queue on the context stream and do not synchronise. ---- */
/* active[t] = 1 when some vertex of t lies strictly within radius of (cx, cy, cz), FP32
 * |p - c|^2 < radius^2 (replaces AdaptiveZone.update_by_proximity, cons/adaptive_zone.py:35-48). */
int mgpbd_zone_proximity(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev, const float* v_p_dev,
                         float cx, float cy, float cz, float radius, int32_t* active_dev);
/* active[t] = 1 when max over the vertices of t of |x - x_old| / dt exceeds threshold, else 0
 * (replaces AdaptiveZone.update_by_strain, cons/adaptive_zone.py:50-64). dt > 0. */
int mgpbd_zone_strain(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i_dev, const float* v_p_dev,
                      const float* v_p_old_dev, float threshold, float dt, int32_t* active_dev);
/* n_rings rings of face neighbours added to the zone, in place; nbr_dev: n_tets x 4 face
 * neighbours, -1 padded (mgpbd_face_adjacency); scratch_dev: n_tets ints, not aliasing active
 * (replaces AdaptiveZone.expand_zone, cons/adaptive_zone.py:66-80). */
int mgpbd_zone_expand(mgpbd_ctx* ctx, int n_tets, const int32_t* nbr_dev, int32_t* active_dev,
                      int32_t* scratch_dev, int n_rings);
/* saved = alpha; alpha[t] = passive_value where active[t] == 0 (replaces
 * MGPBDAdaptiveSolver._mask_inactive_kernel, cons/mgpbd_adaptive.py:64-73, passive_value
 * 1e20 there); the restore is a copy of saved back into alpha. */
int mgpbd_zone_mask_alpha(mgpbd_ctx* ctx, int n_tets, const int32_t* active_dev, float* alpha_dev,
                          float* saved_dev, float passive_value);

/* ---- Learned warm start (extension): a message-passing network over the tet face-adjacency
 * graph (nbr_dev: 4 neighbours per tet, -1 padded, as from mgpbd_face_adjacency) predicting
 * dlambda of the first Newton iteration.
 *   h0 = relu(x W_in + b_in); m_l = mean over face neighbours of h_{l-1};
 *   h_l = relu(h_{l-1} W_l[0:64] + m_l W_l[64:128] + b_l), l = 1..3; y = h3 w_out + b_out;
 *   dlambda = lam_prev + s y  (s = stats[0], the feature scale).
 * Parameters: one FP32 vector of mgpbd_gnn_param_count() values in the order W_in[8][64],
 * b_in[64], {W_l[128][64], b_l[64]} x 3, w_out[64], b_out. The three 128 -> 64 layers run on
 * BF16 MFMA with FP32 accumulation from BF16 weight copies: wt_dev (layer, half, out, in) for
 * the forward pass and w_dev (layer, half, in, out) for the backward pass, 6*64*64 uint16
 * each, written by mgpbd_gnn_pack / mgpbd_gnn_adam. Activations are kept as BF16.
 * ws_dev: mgpbd_gnn_workspace_bytes(n) bytes; mgpbd_gnn_backward reads the activations the
 * preceding mgpbd_gnn_forward left there (same ws_dev, same inputs). */
#define MGPBD_GNN_HIDDEN 64
#define MGPBD_GNN_FEATURES 8
#define MGPBD_GNN_LAYERS 3
int mgpbd_gnn_param_count(void);
int64_t mgpbd_gnn_workspace_bytes(int n);
/* Bytes of the smaller workspace that mgpbd_gnn_features and mgpbd_gnn_infer need (two
 * ping-ponged activation buffers instead of every layer's): a model that is no longer trained
 * keeps only this one. A training workspace serves them as well. */
int64_t mgpbd_gnn_infer_workspace_bytes(int n);
/* Per-tet features x_dev (n x 8): b/A_ii, lam_prev, C, log alpha, log A_ii, rest volume,
 * degree, 1 -- normalised by FP64 statistics written to stats_dev (8 floats: the scale s of
 * b/A_ii and lam_prev, the scale of C, mean/std of log alpha, mean/std of log A_ii, mean
 * volume). lam_prev_dev may be NULL (zero). dinv_dev = 1/A_ii (as cached for the smoother). */
int mgpbd_gnn_features(mgpbd_ctx* ctx, int n, const float* b_dev, const float* dinv_dev,
                       const float* lam_prev_dev, const float* C_dev, const float* alpha_dev,
                       const float* vol_dev, const int32_t* nbr_dev, float* x_dev,
                       float* stats_dev, void* ws_dev);
/* Forward pass; dlam_out_dev (n, may be NULL) = lam_prev + s y. */
int mgpbd_gnn_forward(mgpbd_ctx* ctx, int n, const int32_t* nbr_dev, const float* x_dev,
                      const float* stats_dev, const float* lam_prev_dev,
                      const float* params_dev, const uint16_t* wt_dev, void* ws_dev,
                      float* dlam_out_dev);
/* Inference-only forward (extension): the same network and the same result bit for bit as
 * mgpbd_gnn_forward's dlam_out, with the neighbour mean fused into each MFMA layer (no n x 64
 * mean written or read back), the encoder at 8 channels per thread and the decoder plus
 * prediction computed in the last layer from its tile in LDS. Leaves no activations for a
 * backward pass. */
int mgpbd_gnn_infer(mgpbd_ctx* ctx, int n, const int32_t* nbr_dev, const float* x_dev,
                    const float* stats_dev, const float* lam_prev_dev, const float* params_dev,
                    const uint16_t* wt_dev, void* ws_dev, float* dlam_out_dev);
/* Losses of the forward pass left in ws_dev; each writes dy_dev (n) = d loss / d y and, when
 * loss_dev (device double) is given, the loss.
 *   mse:      mean_i (y_i - (target_i - lam_prev_i) / s)^2
 *   residual: ||A dlam - b||^2 / ||b||^2 with dlam = the forward's dlam_out, A the level-0
 *             SELL-64x4 matrix of the same solve (tmp_dev: 2 n floats). This is the quantity
 *             mgpbd_pcg_solve_guess compares between starts.
 *   energy:   (dlam - t)^T A (dlam - t) / t^T A t with t = target (the converged dlambda): the
 *             A-norm error that bounds CG's convergence from that start (tmp_dev: 2 n).
 *             tat_dev (nullable, one device double) caches t^T A t, which is constant per
 *             training pair: tat_ready 0 computes and stores it, 1 reads it (one SpMV less). */
int mgpbd_gnn_loss_mse(mgpbd_ctx* ctx, int n, const float* stats_dev, const float* lam_prev_dev,
                       const float* target_dev, void* ws_dev, float* dy_dev, double* loss_dev);
int mgpbd_gnn_loss_residual(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr_dev,
                            const int32_t* sell_idx_dev, const float* sell_val_dev,
                            const float* b_dev, const float* dlam_dev, const float* stats_dev,
                            float* tmp_dev, void* ws_dev, float* dy_dev, double* loss_dev);
int mgpbd_gnn_loss_energy(mgpbd_ctx* ctx, int n, const int32_t* slice_ptr_dev,
                          const int32_t* sell_idx_dev, const float* sell_val_dev,
                          const float* target_dev, const float* dlam_dev, const float* stats_dev,
                          float* tmp_dev, void* ws_dev, float* dy_dev, double* loss_dev,
                          double* tat_dev, int tat_ready);
/* Backward pass from dy_dev = d loss / d y: grad_dev gets d loss / d params (FP32,
 * param_count values, overwritten). */
int mgpbd_gnn_backward(mgpbd_ctx* ctx, int n, const int32_t* nbr_dev, const float* x_dev,
                       const float* params_dev, const uint16_t* w_dev, const float* dy_dev,
                       void* ws_dev, float* grad_dev);
/* One Adam step (bias-corrected, step >= 1) on the FP32 parameters; refreshes the BF16 copies
 * when wt_dev and w_dev are given. */
int mgpbd_gnn_adam(mgpbd_ctx* ctx, float* params_dev, const float* grad_dev, float* m_dev,
                   float* v_dev, float lr, float beta1, float beta2, float eps, int step,
                   uint16_t* wt_dev, uint16_t* w_dev);
int mgpbd_gnn_pack(mgpbd_ctx* ctx, const float* params_dev, uint16_t* wt_dev, uint16_t* w_dev);
/* The MFMA building blocks on their own: Y (n x 64 FP32) = act(A1 W1 + A2 W2 + bias) with A*
 * BF16 n x 64 row-major and W*t BF16 [64 out][64 in] (A2/W2t/bias may be NULL, act = relu
 * when relu != 0); and G (64 x 64 FP32) = A^T dZ over n rows (partial_dev:
 * 4096 * ceil(n / 1024) floats, overwritten). */
int mgpbd_gnn_gemm64(mgpbd_ctx* ctx, int n, const uint16_t* A1_dev, const uint16_t* W1t_dev,
                     const uint16_t* A2_dev, const uint16_t* W2t_dev, const float* bias_dev,
                     int relu, float* Y_dev);
int mgpbd_gnn_wgrad64(mgpbd_ctx* ctx, int n, const uint16_t* A_dev, const uint16_t* dZ_dev,
                      float* partial_dev, float* G_dev);

/* ===================================================================================== */
/* Native substep (cons/mgpbd_gpu.py:200-219 and 397-446)                                */
/* ===================================================================================== */

/* Everything one substep reads and writes. Device arrays unless noted; the solver object
 * owns them (the Python MGPBDSolverGPU fills this from its tensors). Sizes: v_p, v_v,
 * v_p_old, f_ext, x_save, dpos 3 n_verts; v_invm n_verts; t_i 4 n_tets; t_c 4 n_cons;
 * Dm_inv 9 n_tets; alpha, lambda, C, rhs, dlam, lam_prev, vol_rows n_cons; gradC 12 n_cons;
 * norms 4 doubles. */
typedef struct mgpbd_substep_desc {
  int32_t n_verts, n_tets, per_tet, n_cons;    /* per_tet 2: ARAP + volume constraint */
  float* v_p; float* v_v; float* v_p_old; const float* v_invm; float* f_ext; /* f_ext: NULL ok */
  const int32_t* t_i; const int32_t* t_c; const float* Dm_inv; const float* alpha;
  float* lambda; float* C; float* gradC; float* rhs; float* dlam; float* lam_prev;
  double* norms; float* x_save; float* dpos;
  /* assembled level 0 (mgpbd_fill_A_shared), its 1/diag and the scatter's incidence */
  const int32_t* slice_ptr; const int32_t* sell_idx; const int32_t* row_len;
  const uint16_t* fill_desc; float* A_val; float* A_dinv; const int32_t* vptr;
  const int32_t* vinc;
  float* pcg_work;                             /* mgpbd_pcg_solve's work buffer */
  /* learned warm start (mgpbd_gnn_features + mgpbd_gnn_infer); NULL without one */
  const int32_t* nbr; float* feat; float* feat_stats; const float* params; const uint16_t* wt;
  void* ws_infer; const float* vol_rows;
  double atol, rtol, tol_cg;                   /* Newton stop and PCG tolerance */
  float gx, gy, gz, dt, damp;
  float drag_target[3]; float drag_k;
  int32_t drag_vertex;                         /* -1: no spring drag */
  int32_t hold_pinned, maxiter, maxiter_cg, relative_tol, use_line_search;
  int32_t spectral_every_newton;               /* 0: spectral refresh at iteration 0 only */
  int32_t coarse_every_newton;                 /* levels 1.. rebuilt: 0 at iteration 0 only, 1 at
                                                  every iteration, k >= 2 at iteration 0 and after
                                                  a solve of >= k PCG iterations */
  int32_t warm_start;                          /* 0 none, 1 previous lambda, 2 learned */
  int32_t overlap_inference;                   /* learned + factored: network on a side stream */
  int32_t factored;                            /* level 0 factored (mgpbd_mg_set_level0_factored) */
} mgpbd_substep_desc;

/* One Newton iteration's record: the reference's (ite, r, cg_iters, omega) plus the PCG's
 * info {||r||^2, ||b||^2, stop reason, ||r0||^2 of the start, start, ||b - A x_guess||^2,
 * ||b - A x_alt||^2, FP64 ||b - A x||^2} (entries 3-6 are -1 for a solve from zero). The
 * iteration that meets the stop test has cg_iters = 0, omega = 1 and no solve (info all -1). */
typedef struct mgpbd_newton_stat {
  int32_t ite, cg_iters;
  double r, omega;
  double info[8];
} mgpbd_newton_stat;

typedef struct mgpbd_substep_state mgpbd_substep_state;
/* Host state of the native substep on one multigrid context: a pinned read-back buffer and,
 * once the learned start overlaps, a side stream (default priority) with its own context. */
mgpbd_substep_state* mgpbd_substep_create(mgpbd_ctx* mg);
void mgpbd_substep_destroy(mgpbd_substep_state* st);
/* One substep in C++: spring drag (drag_vertex >= 0), prediction, the Newton loop of solve()
 * (cons/mgpbd_gpu.py:397-446: lambda = 0; constraint pass; stop when r < atol or
 * r < rtol r0; A fill or the factored level 0; mgpbd_mg_refresh; PCG, from zero or warm
 * started; scatter; the backtracking line search of :379-391 with its full step evaluated by
 * the constraint pass) and the velocity update, on mg's stream, with the same kernels in the
 * same order as the Python loop (bit-identical positions). The host reads one pair of norms
 * per Newton iteration, plus the PCG's polls. stats (host, >= maxiter entries): one record per
 * Newton iteration, *n_stats their count; *predicted = 1 when the network ran.
 * MGPBD_ENONFINITE when ||b||^2 is not finite. */
int mgpbd_substep(mgpbd_substep_state* st, const mgpbd_substep_desc* d, mgpbd_newton_stat* stats,
                  int32_t* n_stats, int32_t* predicted);
/* sizeof(mgpbd_substep_desc) and sizeof(mgpbd_newton_stat), for bindings that restate the
 * structures (ctypes). */
int mgpbd_substep_struct_sizes(int32_t* sizes_out);

#ifdef __cplusplus
}
#endif

#endif /* MGPBD_H */
