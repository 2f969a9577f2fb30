// Native substep: prediction, the MGPBD Newton loop and the velocity update in C++
// (cons/mgpbd_gpu.py:200-219 and 397-446), queued through the library's own entry points in
// the order the Python solver queues them, so positions are bit-identical to its loop. The
// host is only involved where the algorithm needs a host decision: the Newton stop test and
// the line search read one pair of FP64 norms per iteration through a pinned buffer, and the
// PCG polls its done flag as in mgpbd_pcg_solve.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "mgpbd_internal.h"

struct mgpbd_substep_state {
  mgpbd_ctx* mg = nullptr;
  double* host = nullptr;           // pinned: norms[0..2]
  hipStream_t side = nullptr;       // learned start overlapped with the coarse refresh
  mgpbd_ctx* side_ctx = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

#define SUB_TRY(expr)                                                                   \
  do {                                                                                  \
    if (int _e = (expr)) return _e;                                                     \
  } while (0)

// norms[first .. first + count) to the host (synchronises the stream)
int read_norms(mgpbd_substep_state* st, hipStream_t s, const double* norms, int first,
               int count) {
  MGPBD_HIP_TRY(hipMemcpyAsync(st->host + first, norms + first, sizeof(double) * count,
                               hipMemcpyDeviceToHost, s));
  MGPBD_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int constraint_pass(mgpbd_ctx* c, const mgpbd_substep_desc* d) {
  if (d->per_tet == 2)
    return mgpbd_constraints_vol(c, d->n_tets, d->t_i, d->v_p, d->Dm_inv, d->alpha, d->lambda,
                                 d->C, d->gradC, d->rhs, d->norms);
  return mgpbd_constraints(c, d->n_tets, d->t_i, d->v_p, d->Dm_inv, d->alpha, d->lambda, d->C,
                           d->gradC, d->rhs, d->norms);
}

// [current, trial] squared dual norms into norms[1], norms[2] (with_current as in
// mgpbd_dual_norms)
int dual_norms(mgpbd_ctx* c, const mgpbd_substep_desc* d, float omega, int with_current) {
  if (d->per_tet == 2)
    return mgpbd_dual_norms_vol(c, d->n_tets, d->t_i, d->v_p, d->dpos, omega, d->Dm_inv,
                                d->alpha, d->lambda, d->C, with_current, d->norms + 1);
  return mgpbd_dual_norms(c, d->n_tets, d->t_i, d->v_p, d->dpos, omega, d->Dm_inv, d->alpha,
                          d->lambda, d->C, with_current, d->norms + 1);
}

int predict_start(mgpbd_ctx* c, const mgpbd_substep_desc* d) {
  SUB_TRY(mgpbd_gnn_features(c, d->n_cons, d->rhs, d->A_dinv, d->lam_prev, d->C, d->alpha,
                             d->vol_rows, d->nbr, d->feat, d->feat_stats, d->ws_infer));
  return mgpbd_gnn_infer(c, d->n_cons, d->nbr, d->feat, d->feat_stats, d->lam_prev, d->params,
                         d->wt, d->ws_infer, d->dlam);
}

int ensure_side(mgpbd_substep_state* st) {
  if (st->side) return 0;
  MGPBD_HIP_TRY(hipStreamCreateWithFlags(&st->side, hipStreamNonBlocking));
  MGPBD_HIP_TRY(hipEventCreateWithFlags(&st->ev_in, hipEventDisableTiming));
  MGPBD_HIP_TRY(hipEventCreateWithFlags(&st->ev_out, hipEventDisableTiming));
  st->side_ctx = mgpbd_ctx_create(st->side);
  if (!st->side_ctx) return MGPBD_EHIP;
  return 0;
}

// The backtracking line search on the dual residual (cons/mgpbd_gpu.py:379-391): the full
// step is evaluated by the constraint pass at the moved positions, which also leaves the next
// Newton iteration's C, gradC, rhs and ||rhs||^2 when it is taken (*r_next >= 0 then).
int line_search(mgpbd_substep_state* st, mgpbd_ctx* c, hipStream_t s,
                const mgpbd_substep_desc* d, double* omega_out, double* r_next) {
  const int n3 = 3 * d->n_verts;
  *r_next = -1.0;
  SUB_TRY(dual_norms(c, d, 1.0f, 2));                 // current norm into norms[1]
  SUB_TRY(mgpbd_copy(c, n3, d->v_p, d->x_save));
  SUB_TRY(mgpbd_apply_dpos(c, d->n_verts, d->v_p, d->dpos, d->v_invm, 1.0f));
  SUB_TRY(constraint_pass(c, d));                     // trial norm into norms[0]
  SUB_TRY(read_norms(st, s, d->norms, 0, 2));
  const double trial2 = st->host[0], cur2 = st->host[1];
  if (trial2 < cur2) {
    *omega_out = 1.0;
    *r_next = std::sqrt(trial2);
    return 0;
  }
  SUB_TRY(mgpbd_copy(c, n3, d->x_save, d->v_p));
  double omega = 1.0;
  for (int k = 1; k < 10; ++k) {
    omega *= 0.5;
    SUB_TRY(dual_norms(c, d, (float)omega, 0));
    SUB_TRY(read_norms(st, s, d->norms, 2, 1));
    if (st->host[2] < cur2) {
      SUB_TRY(mgpbd_apply_dpos(c, d->n_verts, d->v_p, d->dpos, d->v_invm, (float)omega));
      *omega_out = omega;
      return 0;
    }
  }
  omega *= 0.5;
  SUB_TRY(mgpbd_apply_dpos(c, d->n_verts, d->v_p, d->dpos, d->v_invm, (float)omega));
  *omega_out = omega;
  return 0;
}

int check_desc(const mgpbd_substep_desc* d) {
  if (d->n_verts <= 0 || d->n_tets <= 0 || (d->per_tet != 1 && d->per_tet != 2) ||
      d->n_cons != d->per_tet * d->n_tets || d->maxiter < 0 || d->maxiter_cg < 0 ||
      d->warm_start < 0 || d->warm_start > 2)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: invalid sizes or modes");
  if (!d->v_p || !d->v_v || !d->v_p_old || !d->v_invm || !d->t_i || !d->t_c || !d->Dm_inv ||
      !d->alpha || !d->lambda || !d->C || !d->gradC || !d->rhs || !d->dlam || !d->norms ||
      !d->x_save || !d->dpos || !d->pcg_work || !d->A_dinv)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: null state array");
  if (!d->factored && (!d->slice_ptr || !d->sell_idx || !d->row_len || !d->fill_desc ||
                       !d->A_val || !d->vptr || !d->vinc))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: assembled level 0 needs its pattern");
  if (d->warm_start != 0 && !d->lam_prev)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: warm start needs lam_prev");
  if (d->warm_start == 2 && (!d->nbr || !d->feat || !d->feat_stats || !d->params || !d->wt ||
                             !d->ws_infer || !d->vol_rows))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: learned start needs the network");
  if (d->drag_vertex >= d->n_verts || (d->drag_vertex >= 0 && !d->f_ext))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: drag vertex out of range or no f_ext");
  return 0;
}

void blank(mgpbd_newton_stat* s, int ite, double r) {
  s->ite = ite;
  s->cg_iters = 0;
  s->r = r;
  s->omega = 1.0;
  for (double& v : s->info) v = -1.0;
}

}  // namespace

extern "C" {

mgpbd_substep_state* mgpbd_substep_create(mgpbd_ctx* mg) {
  if (!mg) {
    mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep_create: null context");
    return nullptr;
  }
  auto* st = new mgpbd_substep_state();
  st->mg = mg;
  if (hipHostMalloc((void**)&st->host, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    mgpbd_fail(MGPBD_EHIP, "mgpbd_substep_create: pinned read-back buffer");
    delete st;
    return nullptr;
  }
  return st;
}

void mgpbd_substep_destroy(mgpbd_substep_state* st) {
  if (!st) return;
  if (st->side) hipStreamSynchronize(st->side);
  if (st->side_ctx) mgpbd_ctx_destroy(st->side_ctx);
  if (st->ev_in) hipEventDestroy(st->ev_in);
  if (st->ev_out) hipEventDestroy(st->ev_out);
  if (st->side) hipStreamDestroy(st->side);
  if (st->host) hipHostFree(st->host);
  delete st;
}

int mgpbd_substep(mgpbd_substep_state* st, const mgpbd_substep_desc* d, mgpbd_newton_stat* stats,
                  int32_t* n_stats, int32_t* predicted) {
  if (!st || !d || !stats || !n_stats || !predicted)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep: null argument");
  SUB_TRY(check_desc(d));
  mgpbd_ctx* c = st->mg;
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(c);
  *n_stats = 0;
  *predicted = 0;

  // make_prediction (cons/mgpbd_gpu.py:206-213), with the spring drag before it
  if (d->drag_vertex >= 0)
    SUB_TRY(mgpbd_spring_drag(c, d->n_verts, d->v_p, d->f_ext, d->drag_vertex,
                              d->drag_target[0], d->drag_target[1], d->drag_target[2],
                              d->drag_k));
  SUB_TRY(mgpbd_predict(c, d->n_verts, d->v_p, d->v_v, d->v_p_old, d->v_invm, d->f_ext, d->gx,
                        d->gy, d->gz, d->dt, d->damp, d->hold_pinned));

  // solve (cons/mgpbd_gpu.py:397-446)
  SUB_TRY(mgpbd_fill(c, d->n_cons, 0.0f, d->lambda));
  double r0 = 0.0, r_next = -1.0;
  int prev_cg = 0;   // PCG iterations of the previous Newton iteration's solve
  for (int ite = 0; ite < d->maxiter; ++ite) {
    double r;
    if (r_next < 0.0) {
      SUB_TRY(constraint_pass(c, d));
      SUB_TRY(read_norms(st, s, d->norms, 0, 1));
      const double r2 = st->host[0];
      if (!std::isfinite(r2))
        return mgpbd_fail(MGPBD_ENONFINITE, "mgpbd_substep: non-finite dual residual ||b||^2 = " +
                                                std::to_string(r2));
      r = std::sqrt(r2);
    } else {
      r = r_next;
    }
    r_next = -1.0;
    if (ite == 0) r0 = r;
    mgpbd_newton_stat& rec = stats[(*n_stats)++];
    blank(&rec, ite, r);
    if (r < d->atol || (r0 > 1e-12 && r < d->rtol * r0)) break;

    // level 0 and the coarse levels from the new gradients
    if (!d->factored)
      SUB_TRY(mgpbd_fill_A_shared(c, d->n_cons, d->slice_ptr, d->sell_idx, d->row_len,
                                  d->fill_desc, d->t_c, d->v_invm, d->gradC, d->alpha,
                                  d->A_val, d->A_dinv));
    const int refresh = d->spectral_every_newton || ite == 0;
    // coarse levels: every iteration, the first only, or (k >= 2) also after a long solve
    const int coarse = d->coarse_every_newton == 1 || ite == 0 ||
                       (d->coarse_every_newton >= 2 && prev_cg >= d->coarse_every_newton);
    bool have_guess = false;
    if (ite == 0 && d->warm_start == 2 && d->overlap_inference && d->factored) {
      // the network needs only the constraint pass and level 0's 1/diag: it runs on the
      // side stream while the coarse levels refill on the context's stream
      SUB_TRY(ensure_side(st));
      SUB_TRY(mgpbd_mg_convert_half_levels(c, 0, 1));
      MGPBD_HIP_TRY(hipEventRecord(st->ev_in, s));
      MGPBD_HIP_TRY(hipStreamWaitEvent(st->side, st->ev_in, 0));
      SUB_TRY(predict_start(st->side_ctx, d));
      MGPBD_HIP_TRY(hipEventRecord(st->ev_out, st->side));
      SUB_TRY(mgpbd_mg_refresh(c, refresh, 1, coarse));
      MGPBD_HIP_TRY(hipStreamWaitEvent(s, st->ev_out, 0));
      have_guess = true;
      *predicted = 1;
    } else {
      SUB_TRY(mgpbd_mg_refresh(c, refresh, 0, coarse));
    }

    int it = 0;
    if (ite == 0 && d->warm_start == 1) {
      SUB_TRY(mgpbd_copy(c, d->n_cons, d->lam_prev, d->dlam));
      SUB_TRY(mgpbd_pcg_solve_guess(c, d->rhs, d->dlam, nullptr, d->pcg_work, d->maxiter_cg,
                                    d->tol_cg, d->relative_tol, &it, rec.info));
    } else if (ite == 0 && d->warm_start == 2) {
      if (!have_guess) {
        SUB_TRY(predict_start(c, d));
        *predicted = 1;
      }
      SUB_TRY(mgpbd_pcg_solve_guess(c, d->rhs, d->dlam, d->lam_prev, d->pcg_work,
                                    d->maxiter_cg, d->tol_cg, d->relative_tol, &it, rec.info));
    } else {
      double info[4];
      SUB_TRY(mgpbd_pcg_solve(c, d->rhs, d->dlam, d->pcg_work, d->maxiter_cg, d->tol_cg,
                              d->relative_tol, &it, info));
      rec.info[0] = info[0];
      rec.info[1] = info[1];
      rec.info[2] = info[2];
      rec.info[7] = info[3];
    }
    rec.cg_iters = it;
    prev_cg = it;

    if (d->factored)
      SUB_TRY(mgpbd_mg_factored_scatter(c, d->dlam, d->lambda, d->dpos));
    else
      SUB_TRY(mgpbd_scatter(c, d->n_cons, d->n_verts, d->vptr, d->vinc, d->v_invm, d->gradC,
                            d->dlam, d->lambda, d->dpos));
    if (d->use_line_search) {
      SUB_TRY(line_search(st, c, s, d, &rec.omega, &r_next));
    } else {
      rec.omega = 1.0;
      SUB_TRY(mgpbd_apply_dpos(c, d->n_verts, d->v_p, d->dpos, d->v_invm, 1.0f));
    }
  }
  if (d->warm_start != 0) SUB_TRY(mgpbd_copy(c, d->n_cons, d->lambda, d->lam_prev));

  // update_vel (cons/mgpbd_gpu.py:215-219)
  return mgpbd_update_vel(c, d->n_verts, d->v_p, d->v_p_old, d->v_v, d->dt, d->damp);
}

int mgpbd_substep_struct_sizes(int32_t* sizes_out) {
  if (!sizes_out) return mgpbd_fail(MGPBD_EINVAL, "mgpbd_substep_struct_sizes: null output");
  sizes_out[0] = (int32_t)sizeof(mgpbd_substep_desc);
  sizes_out[1] = (int32_t)sizeof(mgpbd_newton_stat);
  return 0;
}

}  // extern "C"
