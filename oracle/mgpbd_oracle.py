"""ORACLE (test infrastructure only) -- CPU restatement of ``MGPBDSolverGPU``
(cons/mgpbd_gpu.py) and ``GPUMultigridSolver`` (cons/mgpcg_gpu.py).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
use it, and only as the checker / CPU baseline. It is independent of the product package:
structure comes from ``oracle/structure.py`` (literal set/dict restatement) and scipy, the
numerics from ``oracle/liboracle_{f32,f64}.so`` (``oracle/mgpbd_oracle.c``).

The Newton loop below follows ``MGPBDSolverGPU.solve`` (cons/mgpbd_gpu.py:397-446) line by
line, including the quirks the GPU backend keeps in its reference mode: lambda reset per
substep, line search with omega in {1, 1/2, ..., 1/2^10}, pinned vertices moved by the
prediction (make_prediction has no invm guard), the init-time Galerkin pattern.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as sp

from . import structure

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)


def build_oracle() -> None:
    """Compiles liboracle_{f32,f64}.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(precision: str = "f32"):
    if precision not in _LIBS:
        path = os.path.join(_HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build_oracle()
        L = ctypes.CDLL(path)
        R = F32P if precision == "f32" else F64P
        rt = ctypes.c_float if precision == "f32" else ctypes.c_double
        L.orc_rest_shape.argtypes = [ctypes.c_int, I32P, R, R, ctypes.c_double, ctypes.c_double, R, R, R]
        L.orc_predict.argtypes = [ctypes.c_int, R, R, R, R, R, rt, rt, rt, rt, rt, ctypes.c_int]
        L.orc_update_vel.argtypes = [ctypes.c_int, R, R, R, rt, rt]
        L.orc_apply_dpos.argtypes = [ctypes.c_int, R, R, R, rt]
        L.orc_rest_shape_vol.argtypes = [ctypes.c_int, I32P, R, R, ctypes.c_double, R,
                                         ctypes.c_double, ctypes.c_double, R, R, R]
        L.orc_constraints.argtypes = [ctypes.c_int, I32P, R, R, R, R, R, R, R, F64P]
        L.orc_constraints_vol.argtypes = [ctypes.c_int, I32P, R, R, R, R, R, R, R, F64P]
        L.orc_dual_norm2.argtypes = [ctypes.c_int, I32P, R, R, rt, R, R, R]
        L.orc_dual_norm2.restype = ctypes.c_double
        L.orc_dual_norm2_vol.argtypes = [ctypes.c_int, I32P, R, R, rt, R, R, R]
        L.orc_dual_norm2_vol.restype = ctypes.c_double
        L.orc_fill_A.argtypes = [ctypes.c_int, I32P, I32P, I32P, R, R, R, R]
        L.orc_cache_diag_inv.argtypes = [ctypes.c_int, I32P, I32P, R, R]
        L.orc_spmv.argtypes = [ctypes.c_int, I32P, I32P, R, R, R]
        L.orc_jacobi.argtypes = [ctypes.c_int, I32P, I32P, R, R, R, R, R, rt]
        L.orc_refill.argtypes = [ctypes.c_int, I32P, I32P, R, R, R, R]
        L.orc_vcycle.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, rt, R, R]
        L.orc_pcg.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, rt, ctypes.c_int,
                              ctypes.c_double, ctypes.c_int, R, R, R, F64P]
        L.orc_pcg.restype = ctypes.c_int
        L.orc_scatter.argtypes = [ctypes.c_int, ctypes.c_int, I32P, R, R, R, R, R]
        L.orc_pattern.argtypes = [ctypes.c_int, ctypes.c_int, I32P, I32P, I32P]
        L.orc_pattern.restype = ctypes.c_int64
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _LIBS[precision] = (L, R, np.float32 if precision == "f32" else np.float64)
    return _LIBS[precision]


def ptr(a: np.ndarray | None, P):
    if a is None:
        return ctypes.cast(None, P)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P)


def sparsity_pattern(t_i: np.ndarray, n_tets: int, n_verts: int):
    """CSR pattern of A (cons/mgpbd_gpu.py:153-194): the literal set-based restatement for
    small meshes, the C restatement ``orc_pattern`` (same rows) above 20,000 tets."""
    if n_tets <= 20000:
        return structure.build_sparsity_pattern(t_i, n_tets)
    L, _, _ = lib("f32")
    indptr = np.zeros(n_tets + 1, np.int32)
    nnz = L.orc_pattern(n_tets, n_verts, ptr(t_i, I32P), ptr(indptr, I32P), ptr(None, I32P))
    indices = np.zeros(nnz, np.int32)
    L.orc_pattern(n_tets, n_verts, ptr(t_i, I32P), ptr(indptr, I32P), ptr(indices, I32P))
    return indptr, indices


def galerkin_contribution_maps_fast(A_fine: sp.csr_matrix, agg: np.ndarray, Ac: sp.csr_matrix):
    """Vectorised equivalent of ``structure.galerkin_contribution_maps`` for unit
    piecewise-constant P: the contributions of coarse entry (I, J) are the fine CSR positions
    a with (agg[row(a)], agg[col(a)]) == (I, J) in ascending a -- the reference's (fine row of
    I ascending, CSR entry) order, because CSR positions grow with the row. Checked against
    the literal triple loop in tests/test_structure.py."""
    n_f = A_fine.shape[0]
    n_c = Ac.shape[0]
    rows = np.repeat(np.arange(n_f, dtype=np.int64), np.diff(A_fine.indptr))
    key = agg[rows].astype(np.int64) * n_c + agg[A_fine.indices].astype(np.int64)
    ac_rows = np.repeat(np.arange(n_c, dtype=np.int64), np.diff(Ac.indptr))
    ac_key = ac_rows * n_c + Ac.indices.astype(np.int64)   # ascending (sorted CSR)
    pos = np.searchsorted(ac_key, key)
    pos_c = np.minimum(pos, len(ac_key) - 1)
    hit = (pos < len(ac_key)) & (ac_key[pos_c] == key)
    a_idx = np.nonzero(hit)[0]
    order = np.argsort(pos[a_idx], kind="stable")
    fine_idx = a_idx[order].astype(np.int32)
    counts = np.bincount(pos[a_idx], minlength=len(ac_key))
    offsets = np.zeros(len(ac_key) + 1, dtype=np.int32)
    offsets[1:] = np.cumsum(counts)
    return fine_idx, offsets


def factored_level0(t_i, v_invm, gradC, alpha_tilde, A_diag):
    """The scaled level-0 matrix D^-1/2 A D^-1/2 in factored form K K^T + diag(a^), FP64.
    A = J W J^T + diag(alpha~) (cons/mgpbd_gpu.py:273-309: A_ij = sum over shared vertices of
    w_v g_iv . g_jv, A_ii adds alpha~_i), so K_tk = sqrt(w_v(t,k)) g_tk / sqrt(A_tt) and
    a^_t = alpha~_t / A_tt. Returns K (n_tets, 4, 3) and a^ (n_tets,)."""
    tv = np.asarray(t_i, np.int64).reshape(-1, 4)
    isd = 1.0 / np.sqrt(np.asarray(A_diag, np.float64))
    K = (np.sqrt(np.asarray(v_invm, np.float64))[tv][:, :, None] *
         np.asarray(gradC, np.float64).reshape(-1, 4, 3) * isd[:, None, None])
    return K, np.asarray(alpha_tilde, np.float64) * isd * isd


def factored_level0_weighted(t_i, v_invm, gradC, alpha_tilde, A_diag):
    """The same scaled matrix as K W' K^T + diag(a^) with the vertex weights outside K:
    K_tk = c g_tk / sqrt(A_tt), w'_v = w_v / c^2, c^2 = max w (FP64). A tet's constraint is a
    function of its deformation gradient, so its four K_tk sum to zero. Returns K
    (n_tets, 4, 3), a^ (n_tets,) and w' (n_verts,)."""
    tv = np.asarray(t_i, np.int64).reshape(-1, 4)
    w = np.asarray(v_invm, np.float64)
    c2 = w.max() if w.max() > 0 else 1.0
    isd = 1.0 / np.sqrt(np.asarray(A_diag, np.float64))
    K = np.sqrt(c2) * np.asarray(gradC, np.float64).reshape(-1, 4, 3) * isd[:, None, None]
    return K, np.asarray(alpha_tilde, np.float64) * isd * isd, w / c2


def round_storage(x, fmt):
    """float32 values rounded to the V-cycle's 16-bit storage format (round to nearest even;
    FP16 saturates at +-65504), returned as float64."""
    x = np.asarray(x, np.float32)
    if fmt == "fp16":
        return np.clip(x, -65504.0, 65504.0).astype(np.float16).astype(np.float64)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def factored_fourth(K3, fmt):
    """K of a tet's fourth vertex from the stored first three (K3: (..., 3, 3) values of the
    storage format): the negative sum ((k0 + k1) + k2) in float32, rounded to the format."""
    k = np.asarray(K3, np.float32)
    return round_storage(-((k[..., 0, :] + k[..., 1, :]) + k[..., 2, :]), fmt)


def factored_apply(t_i, K, ahat, x, n_verts, w=None):
    """(K W K^T + diag(a^)) x: per-vertex sums u_v = w_v sum_t K_tk x_t (w = 1 when not
    given), then per tet a^_t x_t + sum_k K_tk . u_v(t,k) (FP64)."""
    tv = np.asarray(t_i, np.int64).reshape(-1, 4)
    x = np.asarray(x, np.float64)
    u = np.zeros((n_verts, 3))
    np.add.at(u, tv.reshape(-1), (K * x[:, None, None]).reshape(-1, 3))
    if w is not None:
        u *= np.asarray(w, np.float64)[:, None]
    return ahat * x + np.einsum("tkc,tkc->t", K, u[tv])


class _OrcLevel(ctypes.Structure):
    pass


def _level_struct(R, scalar):
    class OrcLevel(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int),
                    ("ip", I32P), ("ix", I32P), ("a", R), ("dinv", R),
                    ("pip", I32P), ("pix", I32P), ("pa", R),
                    ("tip", I32P), ("tix", I32P), ("ta", R),
                    ("b", R), ("x", R), ("xn", R), ("r", R), ("tmp", R),
                    ("dense", F64P), ("omega", scalar), ("cheb", R)]
    return OrcLevel


class OracleMG:
    """GPUMultigridSolver restated (cons/mgpcg_gpu.py): Galerkin setup with the
    value-dependent pattern, contribution-map refill, V-cycle, PCG."""

    def __init__(self, n_smooth=2, omega_jacobi=2.0 / 3.0, maxiter_cg=100, tol_cg=1e-5,
                 precision="f32", relative=False, structural=False, coarse_solver="jacobi"):
        self.n_smooth = n_smooth
        self.omega = omega_jacobi
        self.maxiter_cg = maxiter_cg
        self.tol_cg = tol_cg
        self.precision = precision
        self.relative = relative
        self.structural = structural
        self.coarse_solver = coarse_solver
        self.L, self.R, self.dt = lib(precision)

    def setup(self, A0: sp.csr_matrix, Ps: list):
        """cons/mgpcg_gpu.py:30-173 (pattern of A0 kept, explicit zeros included)."""
        dt = self.dt
        self.As = [A0]
        self.Ps = [P.tocsr().astype(np.float64) for P in Ps]
        self.contrib = []
        A_prev = A0.astype(np.float64)
        for P in self.Ps:
            if self.structural:
                S = sp.csr_matrix((np.ones_like(A_prev.data), A_prev.indices, A_prev.indptr),
                                  shape=A_prev.shape)
                Ac = (P.T @ S @ P).tocsr()
                Ac.sort_indices()
            else:
                Ac = (P.T @ A_prev @ P).tocsr()
            agg = P.indices.astype(np.int64)
            fi, off = galerkin_contribution_maps_fast(A_prev, agg, Ac)
            self.contrib.append((fi, off))
            self.As.append(Ac)
            A_prev = Ac
        self.levels = []
        for l, A in enumerate(self.As):
            n = A.shape[0]
            lv = {"n": n, "ip": A.indptr.astype(np.int32), "ix": A.indices.astype(np.int32),
                  "a": A.data.astype(dt), "dinv": np.zeros(n, dt)}
            for k in ("b", "x", "xn", "r", "tmp"):
                lv[k] = np.zeros(n, dt)
            if l < len(self.Ps):
                P = self.Ps[l]
                PT = P.T.tocsr()
                lv.update(pip=P.indptr.astype(np.int32), pix=P.indices.astype(np.int32),
                          pa=P.data.astype(dt), tip=PT.indptr.astype(np.int32),
                          tix=PT.indices.astype(np.int32), ta=PT.data.astype(dt))
            self.levels.append(lv)
        for lv in self.levels:
            self.L.orc_cache_diag_inv(lv["n"], ptr(lv["ip"], I32P), ptr(lv["ix"], I32P),
                                      ptr(lv["a"], self.R), ptr(lv["dinv"], self.R))
        self._build_structs()

    def set_level_omegas(self, omegas, cheb=None):
        """Per-level Jacobi weights (None: the common ``omega`` on every level) and, optionally,
        per-level Chebyshev (c1, c2) of smoothing steps 2..17 (arrays of shape (16, 2), or None
        for a Jacobi level)."""
        self.level_omegas = None if omegas is None else [float(w) for w in omegas]
        self.level_cheb = None if cheb is None else [
            None if c is None else np.ascontiguousarray(np.asarray(c).reshape(-1), self.dt)
            for c in cheb]
        self._build_structs()

    def _build_structs(self):
        scalar = ctypes.c_float if self.dt == np.float32 else ctypes.c_double
        S = _level_struct(self.R, scalar)
        oms = getattr(self, "level_omegas", None)
        arr = (S * len(self.levels))()
        R = self.R
        for i, lv in enumerate(self.levels):
            s = arr[i]
            s.n = lv["n"]
            s.ip, s.ix = ptr(lv["ip"], I32P), ptr(lv["ix"], I32P)
            s.a, s.dinv = ptr(lv["a"], R), ptr(lv["dinv"], R)
            for k in ("pip", "pix", "tip", "tix"):
                setattr(s, k, ptr(lv.get(k), I32P))
            for k in ("pa", "ta"):
                setattr(s, k, ptr(lv.get(k), R))
            for k in ("b", "x", "xn", "r", "tmp"):
                setattr(s, k, ptr(lv[k], R))
            s.dense = ptr(lv.get("dense"), F64P)
            s.omega = oms[i] if oms is not None else 0.0
            ch = getattr(self, "level_cheb", None)
            s.cheb = ptr(ch[i] if ch is not None else None, R)
        self._structs = arr

    def update_fine_matrix_data(self, data):
        """cons/mgpcg_gpu.py:175-181."""
        lv = self.levels[0]
        lv["a"][:] = data
        self.L.orc_cache_diag_inv(lv["n"], ptr(lv["ip"], I32P), ptr(lv["ix"], I32P),
                                  ptr(lv["a"], self.R), ptr(lv["dinv"], self.R))

    def update_all_coarse_matrices(self):
        """cons/mgpcg_gpu.py:200-212 (unit p values)."""
        for l, (fi, off) in enumerate(self.contrib):
            f, c = self.levels[l], self.levels[l + 1]
            ones = np.ones(max(len(fi), 1), self.dt)
            self.L.orc_refill(len(off) - 1, ptr(off, I32P), ptr(fi, I32P), ptr(ones, self.R),
                              ptr(ones, self.R), ptr(f["a"], self.R), ptr(c["a"], self.R))
            self.L.orc_cache_diag_inv(c["n"], ptr(c["ip"], I32P), ptr(c["ix"], I32P),
                                      ptr(c["a"], self.R), ptr(c["dinv"], self.R))
        self._factor_coarsest()

    def _factor_coarsest(self):
        """Dense FP64 inverse of the coarsest level (numpy Cholesky), for coarse_solver
        'cholesky' (extension; the reference smooths the coarsest level instead)."""
        if self.coarse_solver != "cholesky" or len(self.levels) < 2:
            return
        c = self.levels[-1]
        A = sp.csr_matrix((c["a"].astype(np.float64), c["ix"], c["ip"]),
                          shape=(c["n"], c["n"])).toarray()
        Lc = np.linalg.cholesky(A)
        Li = np.linalg.inv(Lc)
        inv = Li.T @ Li
        c["dense"] = np.asfortranarray(inv).reshape(-1, order="F").copy()
        self._build_structs()

    def vcycle(self, r):
        z = np.zeros_like(r)
        self.L.orc_vcycle(len(self.levels), ctypes.cast(self._structs, ctypes.c_void_p),
                          self.n_smooth, self.omega, ptr(np.ascontiguousarray(r, self.dt), self.R),
                          ptr(z, self.R))
        return z

    def solve(self, b, x):
        """cons/mgpcg_gpu.py:261-329; returns cg_iters (info in self.last_info)."""
        n = self.levels[0]["n"]
        work = np.zeros(4 * n, self.dt)
        info = np.zeros(3, np.float64)
        it = self.L.orc_pcg(len(self.levels), ctypes.cast(self._structs, ctypes.c_void_p),
                            self.n_smooth, self.omega, self.maxiter_cg, self.tol_cg,
                            int(self.relative), ptr(b, self.R), ptr(x, self.R),
                            ptr(work, self.R), ptr(info, F64P))
        self.last_info = info
        return it


def constraint_vertices(t_i, per_tet: int) -> np.ndarray:
    """Vertex ids of every constraint (4 per constraint, flat int32): the tets' own ids, or
    each tet's ids twice (constraints 2t and 2t+1) with the volume constraint."""
    t = np.ascontiguousarray(t_i, np.int32).reshape(-1, 4)
    return np.ascontiguousarray(np.repeat(t, per_tet, axis=0).reshape(-1))


def expand_prolongation(P, per_tet: int):
    """Level-0 prolongation of the tets' aggregates over their constraints: both constraints
    of a tet go to the tet's aggregate (rows 2t and 2t+1 copy row t)."""
    if per_tet == 1:
        return P
    return P.tocsr()[np.repeat(np.arange(P.shape[0]), per_tet)].tocsr()


class OracleMGPBD:
    """MGPBDSolverGPU restated on the CPU (cons/mgpbd_gpu.py:20-446).

    ``volume_lambda`` (extension; None keeps the reference's single ARAP constraint per tet):
    Lame lambda per tet (array) or for all tets (float); every tet then also carries the volume
    constraint det F - 1 with compliance 1 / (lambda dt^2 vol), as constraint 2t+1 beside its
    ARAP constraint 2t."""

    def __init__(self, v_p, v_p_ref, v_invm, t_i, t_m, gravity, dt, mu=1e6, damp=1.0,
                 maxiter=20, atol=1e-4, rtol=1e-2, maxiter_cg=100, tol_cg=1e-5, n_smooth=2,
                 omega_jacobi=2.0 / 3.0, setup_interval=10000, use_line_search=True,
                 aggregation="tet", benchmark=None, precision="f32", hold_pinned=False,
                 mu_per_tet=None, relative_tol=False, f_ext=None, volume_lambda=None,
                 omega_levels=None):
        self.L, self.R, self.dtype = lib(precision)
        dtp = self.dtype
        self.precision = precision
        self.v_p = np.array(v_p, dtype=dtp).reshape(-1, 3).copy()
        self.v_p_ref = np.array(v_p_ref, dtype=dtp).reshape(-1, 3).copy()
        self.v_invm = np.array(v_invm, dtype=dtp).copy()
        self.t_i = np.array(t_i, dtype=np.int32).reshape(-1).copy()
        self.n_verts = self.v_p.shape[0]
        self.n_tets = self.t_i.shape[0] // 4
        self.gravity = np.array(gravity, dtype=dtp)
        self.dt, self.mu, self.damp = dt, mu, damp
        self.maxiter, self.atol, self.rtol = maxiter, atol, rtol
        self.maxiter_cg, self.tol_cg = maxiter_cg, tol_cg
        self.n_smooth, self.omega_jacobi = n_smooth, omega_jacobi
        self.use_line_search = use_line_search
        self.benchmark = benchmark
        self.hold_pinned = hold_pinned
        self.relative_tol = relative_tol
        self.mu_per_tet = None if mu_per_tet is None else np.asarray(mu_per_tet, dtp).copy()
        self.f_ext = None if f_ext is None else np.asarray(f_ext, dtp).reshape(-1).copy()
        self.volume = volume_lambda is not None
        self.lam_per_tet, self.lam = None, 0.0
        if self.volume:
            if np.ndim(volume_lambda) == 0:
                self.lam = float(volume_lambda)
            else:
                self.lam_per_tet = np.asarray(volume_lambda, dtp).reshape(-1).copy()
        self.per_tet = 2 if self.volume else 1
        self.omega_levels = omega_levels   # fixed Jacobi weight per level (last repeating)
        nt, nv = self.n_tets, self.n_verts
        nc = self.n_cons = self.per_tet * nt
        self.t_c = constraint_vertices(self.t_i, self.per_tet) if self.volume else self.t_i
        self.v_v = np.zeros((nv, 3), dtp)
        self.v_p_old = np.zeros((nv, 3), dtp)
        self.dpos = np.zeros((nv, 3), dtp)
        self.Dm_inv = np.zeros(9 * nt, dtp)
        self.rest_vol = np.zeros(nt, dtp)
        self.alpha_tilde = np.zeros(nc, dtp)
        self.lambdaf = np.zeros(nc, dtp)
        self.constraints = np.zeros(nc, dtp)
        self.gradC = np.zeros(12 * nc, dtp)
        self.b = np.zeros(nc, dtp)
        self.dlam = np.zeros(nc, dtp)
        self.history = []

    def _p(self, a):
        return ptr(a, self.R)

    def init(self, Ps=None, structural=False):
        """cons/mgpbd_gpu.py:92-114. ``Ps`` replaces the reference tet hierarchy (used to give
        the oracle the same extended levels as a GPU run); ``structural`` builds coarse
        patterns from the structure of A instead of its init-time values."""
        L = self.L
        if self.volume:
            L.orc_rest_shape_vol(self.n_tets, ptr(self.t_i, I32P), self._p(self.v_p_ref),
                                 self._p(self.mu_per_tet), float(self.mu),
                                 self._p(self.lam_per_tet), float(self.lam), float(self.dt),
                                 self._p(self.Dm_inv), self._p(self.rest_vol),
                                 self._p(self.alpha_tilde))
        else:
            L.orc_rest_shape(self.n_tets, ptr(self.t_i, I32P), self._p(self.v_p_ref),
                             self._p(self.mu_per_tet), float(self.mu), float(self.dt),
                             self._p(self.Dm_inv), self._p(self.rest_vol),
                             self._p(self.alpha_tilde))
        nc = self.n_cons
        self.indptr, self.indices = sparsity_pattern(self.t_c, nc, self.n_verts)
        self._compute_C_and_gradC()
        data = np.zeros(len(self.indices), self.dtype)
        self._fill_A(data)
        A = sp.csr_matrix((data.astype(np.float32).astype(np.float64), self.indices.copy(),
                           self.indptr.copy()), shape=(nc, nc))
        if Ps is None:
            Ps, sizes = structure.build_tet_hierarchy(self.t_i, self.v_p_ref.astype(np.float64),
                                                      self.n_tets)
            Ps = [expand_prolongation(Ps[0], self.per_tet)] + list(Ps[1:]) if Ps else Ps
            sizes = [nc] + list(sizes[1:])
        else:
            sizes = [nc] + [P.shape[1] for P in Ps]
        self.Ps, self.sizes = Ps, sizes
        self.mg = OracleMG(self.n_smooth, self.omega_jacobi, self.maxiter_cg, self.tol_cg,
                           self.precision, relative=self.relative_tol, structural=structural)
        self.mg.setup(A, Ps)
        if self.omega_levels:
            w = list(self.omega_levels)
            self.mg.set_level_omegas([w[min(l, len(w) - 1)] for l in range(len(self.mg.levels))])
        self.A_data = data

    def _compute_C_and_gradC(self):
        n2 = np.zeros(1, np.float64)
        fn = self.L.orc_constraints_vol if self.volume else self.L.orc_constraints
        fn(self.n_tets, ptr(self.t_i, I32P), self._p(self.v_p),
                               self._p(self.Dm_inv), self._p(self.alpha_tilde),
                               self._p(self.lambdaf), self._p(self.constraints),
                               self._p(self.gradC), self._p(self.b), ptr(n2, F64P))
        return float(np.sqrt(n2[0]))

    def _fill_A(self, data):
        self.L.orc_fill_A(self.n_cons, ptr(self.indptr, I32P), ptr(self.indices, I32P),
                          ptr(self.t_c, I32P), self._p(self.v_invm), self._p(self.gradC),
                          self._p(self.alpha_tilde), self._p(data))

    def step(self, frame=0):
        self.make_prediction()
        self.solve(frame)
        self.update_vel()

    def make_prediction(self):
        g = self.gravity
        self.L.orc_predict(self.n_verts, self._p(self.v_p), self._p(self.v_v),
                           self._p(self.v_p_old), self._p(self.v_invm), self._p(self.f_ext),
                           g[0], g[1], g[2], self.dt, self.damp, int(self.hold_pinned))

    def update_vel(self):
        self.L.orc_update_vel(self.n_verts, self._p(self.v_p), self._p(self.v_p_old),
                              self._p(self.v_v), self.dt, self.damp)

    def _dual_norm(self, omega=None):
        dx = None if omega is None else self.dpos
        fn = self.L.orc_dual_norm2_vol if self.volume else self.L.orc_dual_norm2
        return float(np.sqrt(fn(
            self.n_tets, ptr(self.t_i, I32P), self._p(self.v_p), self._p(dx),
            0.0 if omega is None else omega, self._p(self.Dm_inv), self._p(self.alpha_tilde),
            self._p(self.lambdaf))))

    def _line_search(self, beta=0.5, max_steps=10):
        """cons/mgpbd_gpu.py:379-391."""
        current = self._dual_norm(None)
        omega = 1.0
        for _ in range(max_steps):
            if self._dual_norm(omega) < current:
                return omega
            omega *= beta
        return omega

    def solve(self, frame=0):
        """cons/mgpbd_gpu.py:397-446."""
        if self.benchmark is not None:
            self.benchmark.begin_frame(frame)
        self.lambdaf[:] = 0
        r0 = 0.0
        rec = []
        for ite in range(self.maxiter):
            r = self._compute_C_and_gradC()
            if ite == 0:
                r0 = r
            if r < self.atol or (r0 > 1e-12 and r < self.rtol * r0):
                rec.append((ite, r, 0, 1.0))
                if self.benchmark is not None:
                    self.benchmark.record_iteration(ite, r)
                break
            self._fill_A(self.A_data)
            self.mg.update_fine_matrix_data(self.A_data)
            self.mg.update_all_coarse_matrices()
            self.dlam[:] = 0
            cg_iters = self.mg.solve(self.b, self.dlam)
            self.L.orc_scatter(self.n_cons, self.n_verts, ptr(self.t_c, I32P),
                               self._p(self.v_invm), self._p(self.gradC), self._p(self.dlam),
                               self._p(self.lambdaf), self._p(self.dpos))
            omega = self._line_search() if self.use_line_search else 1.0
            self.L.orc_apply_dpos(self.n_verts, self._p(self.v_p), self._p(self.dpos),
                                  self._p(self.v_invm), omega)
            rec.append((ite, r, cg_iters, omega))
            if self.benchmark is not None:
                self.benchmark.record_iteration(ite, r, cg_iters=cg_iters, omega=omega)
        self.history.append(rec)
        if self.benchmark is not None:
            self.benchmark.end_frame()
