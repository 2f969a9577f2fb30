"""MGPBD global XPBD solver on one MI355X: drop-in for the reference's ``MGPBDSolverGPU``
(cons/mgpbd_gpu.py:20-446) -- same constructor, defaults, methods and attributes.

Every kernel of the substep is libmgpbd (HIP, gfx950); Python runs the Newton loop exactly
as the reference does (cons/mgpbd_gpu.py:397-446), reading back one scalar per Newton
iteration plus the line-search norms.

Reference behaviour kept by default (parity mode):
* lambda is zeroed at the start of every substep and every PCG starts from dlambda = 0;
* ``make_prediction`` moves every vertex, those with ``v_invm == 0`` included
  (cons/mgpbd_gpu.py:206-213 has no invm guard); only the correction skips them;
* tet aggregation gives at most two levels and the coarse Galerkin pattern is fixed by the
  values of A at ``init`` (scipy drops entries that sum to 0.0);
* the PCG stops on the absolute test ||r||^2 < tol_cg^2; the coarsest level gets
  min(50, n) weighted-Jacobi sweeps.

Keyword-only extensions (no reference counterpart; defaults keep the reference's
behaviour): ``hold_pinned``, ``mu_per_tet``, ``relative_tol``, ``hierarchy``
('reference' | 'extended'), ``galerkin`` ('values' | 'structural'), ``max_coarse``, ``max_levels``,
``coarse_solver`` ('jacobi' | 'cholesky'), ``precision`` ('fp32' | 'fp16' | 'bf16'),
``cycle`` ('V' | 'W'), ``overcorrect`` (scale of the coarse correction), ``smoother``
('jacobi' | 'spectral' | 'chebyshev': per-level weights or polynomial from
lambda_max(D^-1 A)), ``verify_residual``
(1: FP64 check of every converged PCG solution, restart when it fails; 2: measure only),
``fp64_iterate`` (FP64 PCG iterate and outer SpMV accumulation), ``power_iters`` (power
steps per refresh of the spectral estimates once settled), ``smoother_safety`` and
``cheb_ratio`` (upper end safety factor over the estimate and interval ratio of the spectral
smoothers), ``stall_rescue`` (re-estimate the spectra and restart a PCG that stops
contracting; spectral smoothers only), ``smooth_schedule`` (smoothing steps per level, the
last entry repeating), ``omega_levels`` (fixed weighted-Jacobi weight per level, the last entry
repeating, instead of the common ``omega_jacobi``), ``f_ext``
(external vertex forces), ``device``,
``warm_start`` ('none' | 'previous' | 'learned') with ``warm_start_model`` (a
``warm_start.LearnedWarmStart``): the first PCG of each substep starts from the previous
substep's lambda, or from the network's prediction with the previous lambda as the fallback;
the device keeps the start with the smaller residual (``start_criterion='residual'``, the
default) or the lower CG energy x.A x / 2 - b.x, i.e. the smaller A-norm error
(``start_criterion='energy'``), and iterates to the same tolerance.
``record_samples`` (a list) collects training pairs from the first Newton iteration.
``spectral_refresh`` ('newton' | 'substep'): the spectral smoothers re-estimate every level's
spectrum, and the dense coarsest solve (coarse_solver='cholesky') re-inverts its level, after
every Newton iteration's refill, or only at a substep's first Newton iteration (the later
iterations of a substep change the matrices little; the kept inverse is SPD, and the verified
PCG and the stall rescue cover a lagging estimate).
``coarse_refresh`` ('newton' | 'substep' | 'adaptive'): the coarse levels 1.. are rebuilt from
level 0 after every Newton iteration's refill (the reference), or only at a substep's first
Newton iteration (extension: later iterations refresh level 0 alone; the V-cycle stays an SPD
preconditioner of level 0's current matrix, and the PCG still solves that matrix to its
tolerance), or ('adaptive') at the first Newton iteration and after every solve that took at
least ``coarse_refresh_iters`` (default 8) PCG iterations: where the matrices move much within a
substep (stiffness contrasts of 1e4, the drag of an inclusion) a lagging hierarchy multiplies the
iteration counts, where they do not the later refreshes are skipped.
``overlap_inference`` (default True; learned start with the factored level 0): the network runs
on a second stream while the coarse levels refill (bit-identical to running it after them).
``level0`` ('assembled' | 'factored'; reduced precision with galerkin='structural'): level 0 is
used as J W J^T + alpha~ in factored form from the tets' gradients (V-cycle sweeps from FP16/BF16
K W' K^T + diag(a^), PCG products and FP64 checks from FP64 sums of the gradients, level 1 from the
aggregates' summed gradients) and is not reassembled per Newton iteration.
``native_substep`` (default False): ``step`` runs prediction, the Newton loop and the velocity
update in the library (``mgpbd_substep``: the same kernels in the same order, positions
bit-identical to the Python loop, the host reading one pair of norms per Newton iteration);
the Python loop runs instead while a ``benchmark`` or ``record_samples`` is attached (their
per-iteration hooks are Python).
``volume_lambda`` (None: the reference's single ARAP constraint per tet; a float or a per-tet
array: the Lame lambda of cons/vbd.py:44-47): every tet also carries the volume constraint
C_H = det F - 1 (the hydrostatic constraint of cons/deform3d.py:81-91) with compliance
1 / (lambda dt^2 vol). Tet t then owns constraints 2t (ARAP) and 2t+1 (volume): the
per-constraint arrays (``alpha_tilde``, ``lambdaf``, ``constraints``, ``gradC``) hold
``n_cons = 2 n_tets`` entries, A has one row per constraint (sibling rows couple through all
four vertices) and both rows of a tet share its aggregate.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import MGPBDError, check, dptr, np_ptr
from .mgpcg_gpu import GPUMultigridSolver
from .runtime import get_context, new_context, to_device
from .sparse_gpu import GPUSparseMatrix, gpu_copy, gpu_fill_zero
from . import tet_aggregation


class _Scalar:
    """Stands in for the reference's 0-d Taichi field (``self.damp[None]``)."""

    def __init__(self, v: float):
        self.value = float(v)

    def __getitem__(self, key):
        return self.value

    def __setitem__(self, key, v):
        self.value = float(v)


class MGPBDSolverGPU:

    def __init__(self, v_p, v_p_ref, v_invm, t_i, t_m, gravity, dt: float, mu: float = 1e6,
                 damp: float = 1.0, maxiter: int = 20, atol: float = 1e-4, rtol: float = 1e-2,
                 maxiter_cg: int = 100, tol_cg: float = 1e-5, n_smooth: int = 2,
                 omega_jacobi: float = 2.0 / 3.0, setup_interval: int = 10000,
                 use_line_search: bool = True, aggregation: str = "tet", benchmark=None, *,
                 hold_pinned: bool = False, mu_per_tet=None, relative_tol: bool = False,
                 hierarchy: str = "reference", galerkin: str = "values",
                 max_coarse: int = 400, max_levels: int | None = None,
                 coarse_sweeps: int = 50, coarse_solver: str = "jacobi",
                 precision: str = "fp32", cycle: str = "V", overcorrect: float = 1.0,
                 smoother: str = "jacobi", verify_residual: int = 0,
                 fp64_iterate: bool = False, power_iters: int = 2,
                 smoother_safety: float = 1.1, cheb_ratio: float = 12.0,
                 stall_rescue: bool = True, smooth_schedule=None, f_ext=None, warm_start: str = "none",
                 start_criterion: str = "residual",
                 warm_start_model=None, record_samples=None, level0: str = "assembled",
                 spectral_refresh: str = "newton", coarse_refresh: str = "newton",
                 coarse_refresh_iters: int = 8, overlap_inference: bool = True,
                 volume_lambda=None, omega_levels=None, native_substep: bool = False,
                 device=None, verbose: bool = True) -> None:
        if aggregation not in ("tet", "pyamg"):
            raise ValueError("aggregation must be 'tet' or 'pyamg'")
        if spectral_refresh not in ("newton", "substep"):
            raise ValueError("spectral_refresh must be 'newton' or 'substep'")
        if coarse_refresh not in ("newton", "substep", "adaptive"):
            raise ValueError("coarse_refresh must be 'newton', 'substep' or 'adaptive'")
        if coarse_refresh == "adaptive" and int(coarse_refresh_iters) < 2:
            raise ValueError("coarse_refresh_iters must be at least 2")
        if level0 not in ("assembled", "factored"):
            raise ValueError("level0 must be 'assembled' or 'factored'")
        if level0 == "factored" and (precision == "fp32" or galerkin != "structural"):
            raise ValueError("level0='factored' needs precision 'fp16' or 'bf16' and "
                             "galerkin='structural'")
        self.ctx = get_context(device)
        dev = self.device = self.ctx.device
        self.v_p = to_device(v_p, torch.float32, dev, (-1, 3))
        self.v_p_ref = to_device(v_p_ref, torch.float32, dev, (-1, 3))
        self.v_invm = to_device(v_invm, torch.float32, dev, (-1,))
        self.t_i = to_device(t_i, torch.int32, dev, (-1,))
        self.t_m = to_device(t_m, torch.float32, dev, (-1,))
        self.n_verts = int(self.v_p.shape[0])
        self.n_tets = int(self.t_i.shape[0] // 4)
        g = gravity.to_numpy() if hasattr(gravity, "to_numpy") else np.asarray(gravity, np.float64)
        self.gravity = tuple(float(x) for x in np.asarray(g).reshape(3))
        self.dt = float(dt)
        self.mu = float(mu)
        self.maxiter = int(maxiter)
        self.atol = float(atol)
        self.rtol = float(rtol)
        self.maxiter_cg = int(maxiter_cg)
        self.tol_cg = float(tol_cg)
        self.n_smooth = int(n_smooth)
        self.omega_jacobi = float(omega_jacobi)
        self.setup_interval = setup_interval      # unused, as in the reference GPU class
        self.use_line_search = bool(use_line_search)
        self.aggregation = aggregation
        self.benchmark = benchmark
        self.damp = _Scalar(damp)
        # extensions
        self.hold_pinned = bool(hold_pinned)
        self.relative_tol = bool(relative_tol)
        self.hierarchy = hierarchy
        self.galerkin = galerkin
        self.max_coarse = int(max_coarse)
        self.max_levels = max_levels
        self.coarse_sweeps = int(coarse_sweeps)
        self.coarse_solver = coarse_solver
        self.precision = precision
        self.cycle = cycle
        self.overcorrect = float(overcorrect)
        self.smoother = smoother
        self.verify_residual = int(verify_residual)
        self.fp64_iterate = bool(fp64_iterate)
        self.power_iters = int(power_iters)
        self.smoother_safety = float(smoother_safety)
        self.cheb_ratio = float(cheb_ratio)
        self.stall_rescue = bool(stall_rescue)
        self.start_criterion = start_criterion
        self.smooth_schedule = smooth_schedule
        self.omega_levels = omega_levels
        self.level0 = level0
        self.spectral_refresh = spectral_refresh
        self.coarse_refresh = coarse_refresh
        self.coarse_refresh_iters = int(coarse_refresh_iters)
        self.overlap_inference = bool(overlap_inference)
        self.native_substep = bool(native_substep)
        self._native = None        # (state handle, descriptor, stats buffer) of mgpbd_substep
        self._side = None          # (stream, context, events) of the overlapped inference
        self._predicted = False
        self.last_predicted = False   # the network ran for this substep's first solve
        self.verbose = verbose
        self.mu_per_tet = None if mu_per_tet is None else to_device(mu_per_tet, torch.float32, dev, (-1,))
        self.volume = volume_lambda is not None
        self.lam_per_tet, self.lam = None, 0.0
        if self.volume:
            if np.ndim(volume_lambda) == 0 and not isinstance(volume_lambda, torch.Tensor):
                self.lam = float(volume_lambda)
                if not self.lam > 0:
                    raise ValueError("volume_lambda must be > 0")
            else:
                self.lam_per_tet = to_device(volume_lambda, torch.float32, dev, (-1,))
                if self.lam_per_tet.numel() != self.n_tets:
                    raise ValueError("volume_lambda needs one value per tet")
        self.per_tet = 2 if self.volume else 1
        self.n_cons = self.per_tet * self.n_tets
        # vertex ids of every constraint (each tet's ids twice with the volume constraint)
        self.t_c = (self.t_i.view(-1, 4).repeat_interleave(2, dim=0).reshape(-1).contiguous()
                    if self.volume else self.t_i)
        self.f_ext = None if f_ext is None else to_device(f_ext, torch.float32, dev, (-1, 3))
        if warm_start not in ("none", "previous", "learned"):
            raise ValueError("warm_start must be 'none', 'previous' or 'learned'")
        if start_criterion not in ("residual", "energy"):
            raise ValueError("start_criterion must be 'residual' or 'energy'")
        if warm_start == "learned" and warm_start_model is None:
            raise ValueError("warm_start='learned' needs warm_start_model")
        if record_samples is not None and warm_start_model is None:
            raise ValueError("record_samples needs warm_start_model (it computes the features)")
        self.warm_start = warm_start
        self.warm_start_model = warm_start_model
        self.record_samples = record_samples

        nv, nt, nc = self.n_verts, self.n_tets, self.n_cons
        f32 = dict(dtype=torch.float32, device=dev)
        self.v_v = torch.zeros(nv, 3, **f32)
        self.v_p_old = torch.zeros(nv, 3, **f32)
        self.v_predict = torch.zeros(nv, 3, **f32)
        self.dpos = torch.zeros(nv, 3, **f32)
        self.Dm_inv = torch.zeros(nt, 3, 3, **f32)
        self.rest_vol = torch.zeros(nt, **f32)
        self.alpha_tilde = torch.zeros(nc, **f32)
        self.lambdaf = torch.zeros(nc, **f32)
        self.constraints = torch.zeros(nc, **f32)
        self.gradC = torch.zeros(nc, 4, 3, **f32)
        self._b_gpu = torch.zeros(nc, **f32)
        self._dlam_gpu = torch.zeros(nc, **f32)
        self._lam_prev = torch.zeros(nc, **f32)
        self._norms = torch.zeros(4, dtype=torch.float64, device=dev)
        self._x_save = torch.zeros(nv, 3, **f32)
        self._gpu_solver = None
        self._amg_Ps = None
        self.last_stats = []
        self.first_solve_info = {}
        self.last_solve_info = []
        self._drag = None
        if warm_start_model is not None and warm_start_model.n != nc:
            raise ValueError("warm_start_model was built for a different constraint count")

    # ------------------------------------------------------------------ setup
    def init(self):
        """Rest shape, pattern, hierarchy, multigrid setup (cons/mgpbd_gpu.py:92-114)."""
        self._precompute_rest_shape()
        # rest volume of every constraint's tet (a network feature)
        self._vol_rows = (self.rest_vol.repeat_interleave(self.per_tet) if self.volume
                          else self.rest_vol)
        self._build_sparsity_pattern()
        self._alpha_tilde_np = self.alpha_tilde.cpu().numpy()
        self._compute_C_and_gradC()
        self._fill_A()
        self._build_hierarchy()
        self._gpu_solver = GPUMultigridSolver(
            n_smooth=self.n_smooth, omega_jacobi=self.omega_jacobi,
            maxiter_cg=self.maxiter_cg, tol_cg=self.tol_cg, relative=self.relative_tol,
            coarse_sweeps=self.coarse_sweeps, galerkin=self.galerkin,
            coarse_solver=self.coarse_solver, precision=self.precision, cycle=self.cycle,
            overcorrect=self.overcorrect, smoother=self.smoother, verify=self.verify_residual,
            fp64_iterate=self.fp64_iterate, power_iters=self.power_iters,
            smoother_safety=self.smoother_safety, cheb_ratio=self.cheb_ratio,
            stall_rescue=self.stall_rescue, smooth_schedule=self.smooth_schedule,
            start_criterion=self.start_criterion,
            omega_levels=self.omega_levels, device=self.device.index)
        self._build_incidence()
        if self.level0 == "factored":
            self._gpu_solver.set_factored_level0(self.t_c, self.v_invm, self.gradC,
                                                 self.alpha_tilde, self._vptr_np, self._vinc_np,
                                                 t_i_host=self._t_c_np)
        self._gpu_solver.setup(self._A, self._aggs)
        if self.verbose:
            print(f"MGPBD-GPU: {self.n_verts} verts, {self.n_tets} tets, {self.n_cons} "
                  f"constraints, {self._nnz} nnz, "
                  f"aggregation={self.aggregation}, levels={self._gpu_solver._level_sizes}")

    def _precompute_rest_shape(self):
        if self.volume:
            self.ctx("mgpbd_rest_shape_vol", self.n_tets, dptr(self.t_i), dptr(self.v_p_ref),
                     dptr(self.mu_per_tet), self.mu, dptr(self.lam_per_tet), self.lam, self.dt,
                     dptr(self.Dm_inv), dptr(self.rest_vol), dptr(self.alpha_tilde))
            return
        self.ctx("mgpbd_rest_shape", self.n_tets, dptr(self.t_i), dptr(self.v_p_ref),
                 dptr(self.mu_per_tet), self.mu, self.dt, dptr(self.Dm_inv),
                 dptr(self.rest_vol), dptr(self.alpha_tilde))

    def _build_sparsity_pattern(self):
        """CSR pattern of A (cons/mgpbd_gpu.py:153-194) in the C++ setup code."""
        lib = _lib.load()
        self._t_i_np = self.t_i.cpu().numpy()
        t = self._t_c_np = self.t_c.cpu().numpy() if self.volume else self._t_i_np
        nc = self.n_cons
        p = lib.mgpbd_pattern_build(nc, self.n_verts, np_ptr(t))
        if not p:
            check(-1, "mgpbd_pattern_build")
        nnz = int(lib.mgpbd_pattern_nnz(p))
        indptr = np.empty(nc + 1, np.int32)
        indices = np.empty(max(nnz, 1), np.int32)
        check(lib.mgpbd_pattern_copy(p, np_ptr(indptr), np_ptr(indices)), "mgpbd_pattern_copy")
        lib.mgpbd_pattern_free(p)
        self._A_indptr, self._A_indices = indptr, indices[:nnz]
        self._nnz = nnz
        self._A = GPUSparseMatrix(nc, nnz, device=self.device.index)
        self._A.set_pattern(indptr, indices[:nnz], diag_csr_pos=indptr[1:] - 1)
        # which vertices each (row, column) pair shares, per SELL slot: built once
        A = self._A
        self._A_desc = torch.zeros(max(A.padded, 1), dtype=torch.int16, device=self.device)
        self.ctx("mgpbd_fill_A_desc", nc, dptr(A.slice_ptr), dptr(A.sell_idx),
                 dptr(A.row_len), dptr(self.t_c), dptr(self._A_desc))

    def _build_hierarchy(self):
        if self.aggregation == "pyamg":
            # cons/mgpbd_gpu.py:124-134: pyamg standard aggregation on the matrix pattern,
            # max_coarse 400, Galerkin patterns from the values at init (as pyamg's products)
            vals = self._A.csr_data() if self.galerkin == "values" else None
            self._aggs = tet_aggregation.build_pyamg_hierarchy(
                self._A_indptr, self._A_indices, vals, max_coarse=400,
                max_levels=self.max_levels or 10)
        elif self.hierarchy == "reference":
            self._aggs = tet_aggregation.reference_aggregates(self._t_i_np, self.n_tets,
                                                              max_coarse=self.max_coarse)
        elif self.hierarchy == "extended":
            self._aggs = tet_aggregation.build_extended_hierarchy(
                self._t_i_np, self.n_tets, max_coarse=self.max_coarse,
                max_levels=(self.max_levels - 1) if self.max_levels else 8)
        else:
            raise ValueError("hierarchy must be 'reference' or 'extended'")
        if self.volume and self.aggregation != "pyamg" and self._aggs:
            # both constraints of a tet go to the tet's aggregate
            agg0, n0 = self._aggs[0]
            self._aggs = [(np.repeat(np.asarray(agg0, np.int32), self.per_tet), n0)] + \
                list(self._aggs[1:])
        self._amg_Ps = None   # built lazily by amg_Ps (scipy objects, for inspection)

    @property
    def amg_Ps(self):
        if self._amg_Ps is None:
            self._amg_Ps = [tet_aggregation.build_prolongation_from_aggregates(a, len(a), n)
                            for a, n in self._aggs]
        return self._amg_Ps

    def _build_incidence(self):
        lib = _lib.load()
        vptr = np.empty(self.n_verts + 1, np.int32)
        vinc = np.empty(max(4 * self.n_cons, 1), np.int32)
        check(lib.mgpbd_vertex_incidence(self.n_cons, self.n_verts, np_ptr(self._t_c_np),
                                         np_ptr(vptr), np_ptr(vinc)), "mgpbd_vertex_incidence")
        self._vptr_np, self._vinc_np = vptr, vinc
        self._vptr = torch.from_numpy(vptr).to(self.device)
        self._vinc = torch.from_numpy(vinc).to(self.device)

    # ------------------------------------------------------------------ time stepping
    def step(self, frame: int = 0):
        """predict -> solve -> velocity update (cons/mgpbd_gpu.py:200-204)."""
        if self.native_substep and self.benchmark is None and self.record_samples is None:
            self._native_step()
            return
        self.make_prediction()
        self.solve(frame)
        self.update_vel()

    def _native_descriptor(self):
        """mgpbd_substep_desc over this solver's tensors (fixed after init)."""
        mg = self._gpu_solver
        A = self._A
        d = _lib.SubstepDesc()
        d.n_verts, d.n_tets, d.per_tet, d.n_cons = (self.n_verts, self.n_tets, self.per_tet,
                                                    self.n_cons)
        for field, t in (("v_p", self.v_p), ("v_v", self.v_v), ("v_p_old", self.v_p_old),
                         ("v_invm", self.v_invm), ("t_i", self.t_i), ("t_c", self.t_c),
                         ("Dm_inv", self.Dm_inv), ("alpha", self.alpha_tilde),
                         ("lam", self.lambdaf), ("C", self.constraints), ("gradC", self.gradC),
                         ("rhs", self._b_gpu), ("dlam", self._dlam_gpu),
                         ("lam_prev", self._lam_prev), ("norms", self._norms),
                         ("x_save", self._x_save), ("dpos", self.dpos),
                         ("slice_ptr", A.slice_ptr), ("sell_idx", A.sell_idx),
                         ("row_len", A.row_len), ("fill_desc", self._A_desc), ("A_val", A.data),
                         ("A_dinv", A.diag_inv), ("vptr", self._vptr), ("vinc", self._vinc),
                         ("pcg_work", mg._pcg_work), ("vol_rows", self._vol_rows)):
            setattr(d, field, dptr(t))
        m = self.warm_start_model
        if self.warm_start == "learned":
            for field, t in (("nbr", m.nbr), ("feat", m.x), ("feat_stats", m.stats),
                             ("params", m.params), ("wt", m.wt), ("ws_infer", m.ws_infer)):
                setattr(d, field, dptr(t))
        d.atol, d.rtol, d.tol_cg = self.atol, self.rtol, mg.tol_cg
        d.gx, d.gy, d.gz = self.gravity
        d.dt = self.dt
        d.hold_pinned, d.maxiter, d.maxiter_cg = int(self.hold_pinned), self.maxiter, mg.maxiter_cg
        d.relative_tol, d.use_line_search = int(mg.relative), int(self.use_line_search)
        d.spectral_every_newton = int(self.spectral_refresh == "newton")
        d.coarse_every_newton = {"newton": 1, "substep": 0}.get(self.coarse_refresh,
                                                                self.coarse_refresh_iters)
        d.warm_start = {"none": 0, "previous": 1, "learned": 2}[self.warm_start]
        d.overlap_inference = int(self.overlap_inference)
        d.factored = int(mg.factored)
        return d

    def _native_step(self):
        """One substep through mgpbd_substep; fills last_stats, last_solve_info,
        first_solve_info and last_predicted as the Python loop does."""
        lib = _lib.load()
        mg = self._gpu_solver
        if self._native is None:
            h = lib.mgpbd_substep_create(mg.ctx.handle)
            if not h:
                check(-1, "mgpbd_substep_create")
            self._native = (h, self._native_descriptor(),
                            (_lib.NewtonStat * max(self.maxiter, 1))())
        h, d, stats = self._native
        d.damp = self.damp[None]
        d.f_ext = dptr(self.f_ext)
        if self._drag is not None:
            v, tx, ty, tz, k = self._drag
            d.drag_vertex, d.drag_k = v, k
            d.drag_target[0], d.drag_target[1], d.drag_target[2] = tx, ty, tz
        else:
            d.drag_vertex = -1
        mg.ctx.bind_stream()
        n, pred = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib.mgpbd_substep(h, ctypes.byref(d), stats, ctypes.byref(n), ctypes.byref(pred)),
              "mgpbd_substep")
        guess = self.warm_start != "none"
        self.last_stats, self.last_solve_info = [], []
        for k in range(n.value):
            s = stats[k]
            self.last_stats.append((s.ite, s.r, s.cg_iters, s.omega))
            i = s.info
            if i[1] < 0:            # the iteration that met the stop test: no solve
                continue
            info = {"rr": i[0], "bb": i[1], "reason": int(i[2])}
            if guess and s.ite == 0:
                info.update(rr0=i[3], start=int(i[4]), rr_guess=i[5], rr_alt=i[6])
            info["rr_true"] = i[7]
            mg.last_info = dict(info)
            info["cg_iters"] = s.cg_iters
            self.last_solve_info.append(info)
        if self.last_solve_info:   # else kept, as the Python loop keeps it
            self.first_solve_info = self.last_solve_info[0]
        self.last_predicted = bool(pred.value)

    def __del__(self):
        nat = getattr(self, "_native", None)
        if nat is not None:
            try:
                _lib.load().mgpbd_substep_destroy(nat[0])
            except Exception:
                pass
            self._native = None

    def device_bytes(self) -> dict:
        """HBM footprint in bytes: per-tet/per-vertex state and the level-0 pattern
        ("state"), the multigrid hierarchy with its context ("hierarchy") and the warm-start
        network when one is attached ("network"); each storage is counted once."""
        from .runtime import device_bytes
        skip = ("_gpu_solver", "warm_start_model", "record_samples")
        seen = set()
        out = {"hierarchy": self._gpu_solver.device_bytes(seen)}
        out["state"] = device_bytes([v for k, v in vars(self).items() if k not in skip], seen=seen)
        if self.warm_start_model is not None:
            out["network"] = device_bytes(self.warm_start_model, seen=seen)
        out["total"] = sum(out.values())
        return out

    def set_drag(self, vertex, target=None, k: float = 1e3):
        """Spring drag loading (extension): before every prediction, f_ext is zero except
        f_ext[vertex] = k (target - x[vertex]); ``vertex=None`` removes it."""
        if vertex is None:
            self._drag = None
            return
        if not 0 <= int(vertex) < self.n_verts:
            raise ValueError("drag vertex out of range")
        if self.f_ext is None:
            self.f_ext = torch.zeros(self.n_verts, 3, dtype=torch.float32, device=self.device)
        t = np.asarray(target, np.float64).reshape(3)
        self._drag = (int(vertex), float(t[0]), float(t[1]), float(t[2]), float(k))

    def make_prediction(self):
        if self._drag is not None:
            v, tx, ty, tz, k = self._drag
            self.ctx("mgpbd_spring_drag", self.n_verts, dptr(self.v_p), dptr(self.f_ext), v,
                     tx, ty, tz, k)
        gx, gy, gz = self.gravity
        self.ctx("mgpbd_predict", self.n_verts, dptr(self.v_p), dptr(self.v_v),
                 dptr(self.v_p_old), dptr(self.v_invm), dptr(self.f_ext), gx, gy, gz,
                 self.dt, self.damp[None], int(self.hold_pinned))

    def update_vel(self):
        self.ctx("mgpbd_update_vel", self.n_verts, dptr(self.v_p), dptr(self.v_p_old),
                 dptr(self.v_v), self.dt, self.damp[None])

    # ------------------------------------------------------------------ kernels
    def _compute_C_and_gradC(self):
        """C, gradC, rhs b = -(C + alpha lambda) and ||b||^2 in one pass."""
        self.ctx("mgpbd_constraints_vol" if self.volume else "mgpbd_constraints", self.n_tets,
                 dptr(self.t_i), dptr(self.v_p),
                 dptr(self.Dm_inv), dptr(self.alpha_tilde), dptr(self.lambdaf),
                 dptr(self.constraints), dptr(self.gradC), dptr(self._b_gpu), dptr(self._norms))

    def _residual_norm(self) -> float:
        r2 = float(self._norms[0].item())
        if not math.isfinite(r2):
            # a NaN dual residual never meets the stop test and would run every Newton and
            # PCG iteration to its cap; report it instead (status MGPBD_ENONFINITE)
            raise MGPBDError(f"non-finite dual residual ||b||^2 = {r2} "
                             f"(status {_lib.MGPBD_ENONFINITE}): NaN/Inf in positions or state")
        return math.sqrt(r2)

    def _fill_A(self):
        A = self._A
        self.ctx("mgpbd_fill_A_shared", self.n_cons, dptr(A.slice_ptr), dptr(A.sell_idx),
                 dptr(A.row_len), dptr(self._A_desc), dptr(self.t_c), dptr(self.v_invm),
                 dptr(self.gradC), dptr(self.alpha_tilde), dptr(A.data), dptr(A.diag_inv))

    def _fill_A_and_update_gpu(self, newton_iter: int = 0, prev_cg: int = 0):
        # the factored level 0 reads the gradients; its assembled values are only wanted for
        # recorded training pairs
        mg = self._gpu_solver
        if not mg.factored or self.record_samples is not None:
            self._fill_A()
        refresh = self.spectral_refresh == "newton" or newton_iter == 0
        coarse = (self.coarse_refresh == "newton" or newton_iter == 0 or
                  (self.coarse_refresh == "adaptive" and prev_cg >= self.coarse_refresh_iters))
        hook = None
        if (newton_iter == 0 and self.warm_start == "learned" and self.overlap_inference and
                mg.factored):
            hook = self._launch_prediction
        mg.update_all_coarse_matrices(refresh_smoother=refresh, after_level0=hook, coarse=coarse)
        if hook is not None:
            torch.cuda.current_stream(self.device).wait_event(self._side[3])

    def _launch_prediction(self):
        """The network's start for the first solve on a second stream (default priority), once
        level 0's 1/diag is queued: it needs only the constraint pass's outputs and that
        diagonal, so it runs while the coarse levels refill on the caller's stream."""
        if self._side is None:
            stream = torch.cuda.Stream(self.device)
            with torch.cuda.stream(stream):
                ctx = new_context(self.device.index)   # bound to the side stream
            self._side = (stream, ctx, torch.cuda.Event(), torch.cuda.Event())
        stream, ctx, ev_in, ev_out = self._side
        ev_in.record(torch.cuda.current_stream(self.device))
        stream.wait_event(ev_in)
        self.warm_start_model.predict(self._b_gpu, self._A.diag_inv, self._lam_prev,
                                      self.constraints, self.alpha_tilde, self._vol_rows,
                                      self._dlam_gpu, ctx=ctx)
        ev_out.record(stream)
        self._predicted = True
        self.last_predicted = True

    def _scatter_dlam_gpu(self):
        mg = self._gpu_solver
        if mg is not None and mg.factored:   # vertex-order gradients of the factored level 0
            mg.ctx("mgpbd_mg_factored_scatter", dptr(self._dlam_gpu), dptr(self.lambdaf),
                   dptr(self.dpos))
            return
        self.ctx("mgpbd_scatter", self.n_cons, self.n_verts, dptr(self._vptr), dptr(self._vinc),
                 dptr(self.v_invm), dptr(self.gradC), dptr(self._dlam_gpu), dptr(self.lambdaf),
                 dptr(self.dpos))

    def _apply_dpos(self, omega: float):
        self.ctx("mgpbd_apply_dpos", self.n_verts, dptr(self.v_p), dptr(self.dpos),
                 dptr(self.v_invm), float(omega))

    def _dual_norms(self, omega: float, with_current: int, read: bool = True):
        """[current, trial] squared dual norms into _norms[1:3] (with_current: 0 trial only,
        1 both, 2 current only); read=False leaves them on the device."""
        self.ctx("mgpbd_dual_norms_vol" if self.volume else "mgpbd_dual_norms", self.n_tets,
                 dptr(self.t_i), dptr(self.v_p),
                 dptr(self.dpos), float(omega), dptr(self.Dm_inv), dptr(self.alpha_tilde),
                 dptr(self.lambdaf), dptr(self.constraints), int(with_current),
                 dptr(self._norms[1:3]))
        return self._norms[1:3].tolist() if read else None

    def _line_search(self, beta=0.5, max_steps=10):
        """Backtracking on the dual residual norm (cons/mgpbd_gpu.py:379-391): trials at
        omega = 1, 1/2, ..., 1/2^9 against the current norm (from the C stored for this Newton
        iterate), 1/2^10 when none is smaller. Applies the step and returns (omega, r): the full
        step is tried by moving the positions and running the constraint pass there, whose
        ||b|| is the trial norm and, when the step is taken, the next Newton iteration's C,
        gradC, b and residual r (None when a shorter step was taken: the positions were
        restored, the shorter step applied, and the next iteration evaluates them)."""
        self._dual_norms(1.0, 2, read=False)           # current norm into _norms[1]
        gpu_copy(self.v_p.view(-1), self._x_save.view(-1), 3 * self.n_verts)
        self._apply_dpos(1.0)
        self._compute_C_and_gradC()
        trial2, cur2 = self._norms[:2].tolist()
        if trial2 < cur2:
            return 1.0, math.sqrt(trial2)
        gpu_copy(self._x_save.view(-1), self.v_p.view(-1), 3 * self.n_verts)
        omega = 1.0
        for k in range(1, max_steps):
            omega *= beta
            _, trial2 = self._dual_norms(omega, 0)
            if trial2 < cur2:
                self._apply_dpos(omega)
                return omega, None
        omega *= beta
        self._apply_dpos(omega)
        return omega, None

    # ------------------------------------------------------------------ Newton loop
    def solve(self, frame: int = 0):
        """MGPBD outer loop (cons/mgpbd_gpu.py:397-446)."""
        if self.benchmark is not None:
            self.benchmark.begin_frame(frame)
        self.last_predicted = False
        self.lambdaf.zero_()
        r0 = 0.0
        stats = []
        self.last_solve_info = []
        r_next = None   # residual of the positions the line search's full step left evaluated
        cg_iters = 0    # the previous Newton iteration's PCG iterations
        for ite in range(self.maxiter):
            if r_next is None:
                self._compute_C_and_gradC()
                r = self._residual_norm()
            else:
                r = r_next
            r_next = None
            if ite == 0:
                r0 = r
            if r < self.atol or (r0 > 1e-12 and r < self.rtol * r0):
                stats.append((ite, r, 0, 1.0))
                if self.benchmark is not None:
                    self.benchmark.record_iteration(ite, r)
                break
            self._fill_A_and_update_gpu(ite, cg_iters)
            if ite == 0 and self.warm_start != "none":
                cg_iters = self._solve_warm_started()
            else:
                cg_iters = self._gpu_solver.solve(self._b_gpu, self._dlam_gpu)
            if ite == 0 and self.record_samples is not None:
                self.record_samples.append(self.warm_start_model.make_sample(
                    self._b_gpu, self._A.diag_inv, self._lam_prev, self.constraints,
                    self.alpha_tilde, self._vol_rows, self._dlam_gpu, self._A))
            self.last_solve_info.append(dict(self._gpu_solver.last_info, cg_iters=cg_iters))
            if ite == 0:
                self.first_solve_info = self.last_solve_info[0]
            self._scatter_dlam_gpu()
            if self.use_line_search:
                omega, r_next = self._line_search()
            else:
                omega = 1.0
                self._apply_dpos(omega)
            stats.append((ite, r, cg_iters, omega))
            if self.benchmark is not None:
                self.benchmark.record_iteration(ite, r, cg_iters=cg_iters, omega=omega)
        self.last_stats = stats
        if self.warm_start != "none" or self.record_samples is not None:
            gpu_copy(self.lambdaf, self._lam_prev, self.n_cons)
        if self.benchmark is not None:
            self.benchmark.end_frame()

    def _solve_warm_started(self) -> int:
        """First PCG of the substep from the previous substep's lambda, or from the learned
        prediction with the previous lambda as the alternative start."""
        if self.warm_start == "previous":
            gpu_copy(self._lam_prev, self._dlam_gpu, self.n_cons)
            return self._gpu_solver.solve_guess(self._b_gpu, self._dlam_gpu)
        if not self._predicted:   # else already queued beside the hierarchy refresh
            self.warm_start_model.predict(self._b_gpu, self._A.diag_inv, self._lam_prev,
                                          self.constraints, self.alpha_tilde, self._vol_rows,
                                          self._dlam_gpu)
            self.last_predicted = True
        self._predicted = False
        return self._gpu_solver.solve_guess(self._b_gpu, self._dlam_gpu, self._lam_prev)
