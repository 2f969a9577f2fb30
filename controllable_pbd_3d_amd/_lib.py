"""ctypes binding of libmgpbd.so (declarations in include/mgpbd.h).

The library is the product: when it is missing the package raises instead of falling back
to any CPU path. Device pointers come from torch tensors on the current HIP device; torch is
used for device memory and streams only.
"""

from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmgpbd.so")

c_int, c_float, c_double, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
c_int64 = ctypes.c_int64
I32P = ctypes.POINTER(ctypes.c_int32)
F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); device pointers are passed as c_void_p
_V = c_void_p
SIGNATURES = {
    "mgpbd_last_error": (ctypes.c_char_p, []),
    "mgpbd_abi_version": (c_int, []),
    # host setup
    "mgpbd_pattern_build": (_V, [c_int, c_int, I32P]),
    "mgpbd_pattern_nnz": (c_int64, [_V]),
    "mgpbd_pattern_copy": (c_int, [_V, I32P, I32P]),
    "mgpbd_pattern_free": (None, [_V]),
    "mgpbd_face_adjacency": (c_int, [c_int, I32P, I32P, I32P]),
    "mgpbd_tet_aggregate": (c_int, [c_int, I32P, I32P, c_int, I32P, I32P]),
    "mgpbd_graph_aggregate": (c_int, [c_int, I32P, I32P, c_int, I32P, I32P]),
    "mgpbd_standard_aggregate": (c_int, [c_int, I32P, I32P, I32P, I32P]),
    "mgpbd_graph_merge_singletons": (c_int, [c_int, I32P, I32P, I32P, I32P]),
    "mgpbd_graph_coarsen": (_V, [c_int, I32P, I32P, I32P, c_int]),
    "mgpbd_graph_nnz": (c_int64, [_V]),
    "mgpbd_graph_copy": (c_int, [_V, I32P, I32P]),
    "mgpbd_graph_free": (None, [_V]),
    "mgpbd_galerkin_build": (_V, [c_int, I32P, I32P, F32P, I32P, c_int]),
    "mgpbd_galerkin_sizes": (c_int, [_V, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "mgpbd_galerkin_copy": (c_int, [_V, I32P, I32P, I32P, I32P, I32P]),
    "mgpbd_galerkin_free": (None, [_V]),
    "mgpbd_vertex_incidence": (c_int, [c_int, c_int, I32P, I32P, I32P]),
    "mgpbd_sell_layout": (c_int, [c_int, I32P, I32P]),
    "mgpbd_sell_map": (c_int, [c_int, I32P, I32P, I32P]),
    "mgpbd_sell_columns": (c_int, [c_int, I32P, I32P, I32P, I32P]),
    "mgpbd_sell_window": (c_int, [c_int, I32P, I32P, I32P, I32P, c_int64]),
    # context
    "mgpbd_ctx_create": (_V, [_V]),
    "mgpbd_ctx_set_stream": (c_int, [_V, _V]),
    "mgpbd_ctx_destroy": (None, [_V]),
    "mgpbd_ctx_sync": (c_int, [_V]),
    "mgpbd_ctx_device_bytes": (c_int, [_V, ctypes.POINTER(c_int64)]),
    "mgpbd_ctx_profile": (c_int, [_V, c_int]),
    "mgpbd_ctx_use_graphs": (c_int, [_V, c_int, c_int]),
    "mgpbd_pcg_set_verify": (c_int, [_V, c_int]),
    "mgpbd_pcg_set_fp64_iterate": (c_int, [_V, c_int]),
    "mgpbd_pcg_set_rescue": (c_int, [_V, c_int, c_float, c_double]),
    "mgpbd_pcg_counters": (c_int, [_V, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "mgpbd_pcg_lag_rescues": (c_int, [_V, ctypes.POINTER(c_int64)]),
    "mgpbd_pcg_set_lag_rescue": (c_int, [_V, c_int]),
    "mgpbd_pcg_set_start_criterion": (c_int, [_V, c_int]),
    "mgpbd_ctx_profile_read": (c_int, [_V, F64P, ctypes.POINTER(c_int64)]),
    # per-tet / per-vertex
    "mgpbd_rest_shape": (c_int, [_V, c_int, _V, _V, _V, c_double, c_double, _V, _V, _V]),
    "mgpbd_rest_shape_vol": (c_int, [_V, c_int, _V, _V, _V, c_double, _V, c_double, c_double,
                                     _V, _V, _V]),
    "mgpbd_constraints_vol": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_dual_norms_vol": (c_int, [_V, c_int, _V, _V, _V, c_float, _V, _V, _V, _V, c_int, _V]),
    "mgpbd_predict": (c_int, [_V, c_int, _V, _V, _V, _V, _V, c_float, c_float, c_float,
                              c_float, c_float, c_int]),
    "mgpbd_update_vel": (c_int, [_V, c_int, _V, _V, _V, c_float, c_float]),
    "mgpbd_constraints": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_constraints_noalloc": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_fill_A": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_fill_A_desc": (c_int, [_V, c_int, _V, _V, _V, _V, _V]),
    "mgpbd_fill_A_shared": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_scatter": (c_int, [_V, c_int, c_int, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_dual_norms": (c_int, [_V, c_int, _V, _V, _V, c_float, _V, _V, _V, _V, c_int, _V]),
    "mgpbd_apply_dpos": (c_int, [_V, c_int, _V, _V, _V, c_float]),
    # sparse / BLAS-1
    "mgpbd_spmv": (c_int, [_V, c_int, _V, _V, _V, _V, _V]),
    "mgpbd_jacobi": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, c_float]),
    "mgpbd_diag_inv": (c_int, [_V, c_int, _V, _V, _V]),
    "mgpbd_coarse_refill": (c_int, [_V, c_int, _V, _V, _V, _V, _V]),
    "mgpbd_restrict": (c_int, [_V, c_int, _V, _V, _V, _V]),
    "mgpbd_prolong": (c_int, [_V, c_int, _V, _V, _V]),
    "mgpbd_dot": (c_int, [_V, c_int, _V, _V, _V]),
    "mgpbd_axpy": (c_int, [_V, c_int, c_float, _V, _V]),
    "mgpbd_xpay": (c_int, [_V, c_int, c_float, _V, _V]),
    "mgpbd_copy": (c_int, [_V, c_int, _V, _V]),
    "mgpbd_fill": (c_int, [_V, c_int, c_float, _V]),
    "mgpbd_scale": (c_int, [_V, c_int, c_float, _V]),
    "mgpbd_negate_and_add": (c_int, [_V, c_int, _V, _V]),
    # multigrid / PCG
    "mgpbd_mg_configure": (c_int, [_V, c_int, c_int, c_float, c_int, c_int]),
    "mgpbd_mg_set_level": (c_int, [_V, c_int, c_int, _V, _V, _V, _V, _V, _V, _V, _V, c_int]),
    "mgpbd_mg_factor_coarsest": (c_int, [_V]),
    "mgpbd_mg_set_coarse_update": (c_int, [_V, c_int]),
    "mgpbd_mg_coarse_stats": (c_int, [_V, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "mgpbd_mg_set_precision": (c_int, [_V, c_int, c_int]),
    "mgpbd_mg_set_level_half": (c_int, [_V, c_int, _V, _V, _V, _V, _V]),
    "mgpbd_mg_set_level_window": (c_int, [_V, c_int, _V, _V, c_int, c_int, _V]),
    "mgpbd_mg_convert_half": (c_int, [_V]),
    "mgpbd_mg_convert_half_levels": (c_int, [_V, c_int, c_int]),
    "mgpbd_mg_set_level0_factored": (c_int, [_V, c_int] + [_V] * 14),
    "mgpbd_mg_set_level1_factored": (c_int, [_V, c_int, _V, _V, _V, _V, c_int, _V, _V, _V, _V, _V, _V]),
    "mgpbd_mg_set_level0_offsets": (c_int, [_V, _V, c_int]),
    "mgpbd_mg_set_level0_packed_rows": (c_int, [_V, _V, c_int, c_int]),
    "mgpbd_mg_factored_refill": (c_int, [_V]),
    "mgpbd_mg_factored_scatter": (c_int, [_V, _V, _V, _V]),
    "mgpbd_mg_level_residual": (c_int, [_V, c_int, _V, _V, _V]),
    "mgpbd_mg_vcycle": (c_int, [_V, _V, _V]),
    "mgpbd_mg_probe_level0": (c_int, [_V, c_int, _V, _V, _V, _V, c_int64, F64P]),
    "mgpbd_mg_set_cycle": (c_int, [_V, c_int, c_float]),
    "mgpbd_mg_set_launch_fusion": (c_int, [_V, c_int]),
    "mgpbd_mg_set_level_smoothing": (c_int, [_V, c_int, c_int]),
    "mgpbd_mg_set_smoother": (c_int, [_V, c_int, c_float, c_float, c_int, c_float]),
    "mgpbd_mg_update_smoother": (c_int, [_V, F64P]),
    "mgpbd_mg_set_level_weights": (c_int, [_V, c_int, F32P]),
    "mgpbd_mg_get_smoother": (c_int, [_V, F32P, F32P]),
    "mgpbd_pcg_solve": (c_int, [_V, _V, _V, _V, c_int, c_double, c_int,
                                ctypes.POINTER(c_int), F64P]),
    # warm start
    "mgpbd_pcg_solve_guess": (c_int, [_V, _V, _V, _V, _V, c_int, c_double, c_int,
                                      ctypes.POINTER(c_int), F64P]),
    "mgpbd_spring_drag": (c_int, [_V, c_int, _V, _V, c_int, c_float, c_float, c_float, c_float]),
    # spatial adaptivity
    "mgpbd_zone_proximity": (c_int, [_V, c_int, _V, _V, c_float, c_float, c_float, c_float, _V]),
    "mgpbd_zone_strain": (c_int, [_V, c_int, _V, _V, _V, c_float, c_float, _V]),
    "mgpbd_zone_expand": (c_int, [_V, c_int, _V, _V, _V, c_int]),
    "mgpbd_zone_mask_alpha": (c_int, [_V, c_int, _V, _V, _V, c_float]),
    # learned warm start
    "mgpbd_gnn_param_count": (c_int, []),
    "mgpbd_gnn_workspace_bytes": (c_int64, [c_int]),
    "mgpbd_gnn_infer_workspace_bytes": (c_int64, [c_int]),
    "mgpbd_gnn_features": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_forward": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_infer": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_loss_mse": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_loss_residual": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_loss_energy": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V,
                                      c_int]),
    "mgpbd_gnn_backward": (c_int, [_V, c_int, _V, _V, _V, _V, _V, _V, _V]),
    "mgpbd_gnn_adam": (c_int, [_V, _V, _V, _V, _V, c_float, c_float, c_float, c_float, c_int,
                               _V, _V]),
    "mgpbd_gnn_pack": (c_int, [_V, _V, _V, _V]),
    "mgpbd_gnn_gemm64": (c_int, [_V, c_int, _V, _V, _V, _V, _V, c_int, _V]),
    "mgpbd_gnn_wgrad64": (c_int, [_V, c_int, _V, _V, _V, _V]),
    # native substep
    "mgpbd_mg_set_refill": (c_int, [_V, c_int, c_int, _V, _V, _V, _V, _V, c_int, _V, _V]),
    "mgpbd_mg_refresh": (c_int, [_V, c_int, c_int, c_int]),
    "mgpbd_mg_set_coarse_lagged": (c_int, [_V, c_int]),
    "mgpbd_substep_struct_sizes": (c_int, [I32P]),
    "mgpbd_substep_create": (_V, [_V]),
    "mgpbd_substep_destroy": (None, [_V]),
    "mgpbd_substep": (c_int, [_V, _V, _V, ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int32)]),
}


class SubstepDesc(ctypes.Structure):
    """mgpbd_substep_desc (include/mgpbd.h): device pointers as addresses."""
    _fields_ = [(n, ctypes.c_int32) for n in ("n_verts", "n_tets", "per_tet", "n_cons")] + \
        [(n, _V) for n in ("v_p", "v_v", "v_p_old", "v_invm", "f_ext", "t_i", "t_c", "Dm_inv",
                           "alpha", "lam", "C", "gradC", "rhs", "dlam", "lam_prev", "norms",
                           "x_save", "dpos", "slice_ptr", "sell_idx", "row_len", "fill_desc",
                           "A_val", "A_dinv", "vptr", "vinc", "pcg_work", "nbr", "feat",
                           "feat_stats", "params", "wt", "ws_infer", "vol_rows")] + \
        [(n, c_double) for n in ("atol", "rtol", "tol_cg")] + \
        [(n, c_float) for n in ("gx", "gy", "gz", "dt", "damp")] + \
        [("drag_target", c_float * 3), ("drag_k", c_float)] + \
        [(n, ctypes.c_int32) for n in ("drag_vertex", "hold_pinned", "maxiter", "maxiter_cg",
                                       "relative_tol", "use_line_search",
                                       "spectral_every_newton", "coarse_every_newton",
                                       "warm_start",
                                       "overlap_inference", "factored")]


class NewtonStat(ctypes.Structure):
    """mgpbd_newton_stat (include/mgpbd.h)."""
    _fields_ = [("ite", ctypes.c_int32), ("cg_iters", ctypes.c_int32), ("r", c_double),
                ("omega", c_double), ("info", c_double * 8)]

_lib = None


MGPBD_EINVAL, MGPBD_EHIP, MGPBD_ENONFINITE, MGPBD_EINTERNAL, MGPBD_ENOMEM = -1, -2, -3, -4, -5


class MGPBDError(RuntimeError):
    """A libmgpbd entry point failed; ``status`` is its MGPBD_E* code when one was returned."""

    def __init__(self, msg: str, status: int | None = None):
        super().__init__(msg)
        self.status = status


def load():
    """Loads libmgpbd.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MGPBDError(f"{LIB_PATH} is missing: build it with "
                         "`python -m controllable_pbd_3d_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = load().mgpbd_last_error().decode(errors="replace")
        raise MGPBDError(f"{what or 'libmgpbd'} failed ({status}): {msg}", status)


def np_ptr(a, P=I32P):
    return a.ctypes.data_as(P)


def dptr(t) -> int | None:
    """Device address of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()
