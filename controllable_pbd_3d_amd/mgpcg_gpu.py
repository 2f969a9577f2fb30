"""Multigrid-preconditioned CG on the GPU, same interface as the reference's
``GPUMultigridSolver`` (cons/mgpcg_gpu.py:18-329).

Differences in mechanism, not in results:
* coarse patterns and refill maps come from libmgpbd's O(nnz) C++ Galerkin builder, which
  reproduces scipy's ``P^T A P`` pattern (including the entries it drops because they sum to
  0.0 at setup) and the reference's contribution order;
* the PCG keeps alpha, beta and the stopping test on the device (FP64 reductions) and the
  host polls a done flag every few iterations instead of syncing three times per iteration;
* Jacobi ping-pongs two buffers instead of copying x_new back to x.

Extensions (keyword-only, no reference counterpart):
* ``coarse_solver='cholesky'``: the coarsest level is solved exactly with an FP64 dense
  inverse (refactored on every coarse refill) instead of min(50, n) sweeps;
* ``precision='fp16' | 'bf16'``: the V-cycle runs on the diagonally scaled levels stored in
  half precision (FP32 accumulation), inside a flexible (Polak-Ribiere) FP32 PCG with FP64
  dot products; the outer SpMV stays FP32;
* ``smoother='spectral'``: per-level Jacobi weights target / (safety lambda_max(D^-1 A)) from a
  device power iteration refreshed at every refill (the reference's fixed 2/3 makes the V-cycle
  indefinite once lambda_max(D^-1 A) > 3, which per-tet stiffness contrast produces);
  ``smoother='chebyshev'``: the same estimates drive a Chebyshev polynomial smoother on
  [lambda_max / cheb_ratio, lambda_max] (same sweeps per phase, fewer PCG iterations);
* ``cycle='W'`` and ``overcorrect``: W-cycle and a scaled coarse correction;
* ``verify=1`` (or True): a converged solve is checked with an FP64 residual of the returned x
  and restarted from x when it misses the tolerance; ``verify=2`` only measures that residual;
* ``fp64_iterate=True``: the PCG iterate is accumulated in FP64 and the outer SpMV takes FP64
  products and sums (on stiff high-contrast systems the rounding of an FP32 iterate alone keeps
  the true residual above 1e-6);
* ``stall_rescue`` (default on with the spectral smoothers): a PCG that stops contracting
  re-estimates the spectra from a fresh vector with a raised safety factor and restarts;
* ``set_factored_level0`` (before ``setup``, reduced precision and the structural Galerkin
  pattern only): level 0 is used in factored form A = J W J^T + diag(alpha~) from the tets'
  gradients and never reassembled: the V-cycle applies K W' K^T + diag(a^) (K_tk = c g_tk /
  sqrt(A_tt), W' = diag(invm) / c^2, c^2 = max invm, stored FP16/BF16 with a tet's fourth
  vertex's K the negative sum of the other three; ~12 bytes per vertex-tet incidence and ~42
  per tet instead of ~4 per matrix entry), the PCG and its FP64 check apply J W J^T + alpha~ with FP64 sums of the
  FP32 gradients (~180 MB per product at 1M tets instead of ~450 MB), and level 1 is
  G W G^T + P^T diag(alpha~) P with G the aggregates' summed gradients per vertex.
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import check, dptr, np_ptr
from .runtime import new_context
from .sparse_gpu import GPUSparseMatrix


def _members(agg: np.ndarray, n_c: int):
    order = np.argsort(agg, kind="stable").astype(np.int32)
    ptr = np.zeros(n_c + 1, np.int32)
    ptr[1:] = np.cumsum(np.bincount(agg, minlength=n_c))
    return ptr, order


def uniform_rows(vptr: np.ndarray, max_growth: float = 1.25) -> int:
    """Rows of four slots per vertex for a uniform vertex-slice layout (every slice as tall as
    the busiest vertex), or 0 when that would exceed max_growth times the per-slice layout's
    slots (or more than 8 rows). Kuhn-cube meshes (at most 24 tets per vertex) give 6."""
    cnt = np.diff(np.asarray(vptr, np.int64))
    nv = len(cnt)
    if nv == 0:
        return 0
    ns = (nv + 15) // 16
    padded = np.zeros(16 * ns, np.int64)
    padded[:nv] = cnt
    per_slice = int((64 * ((padded.reshape(ns, 16).max(axis=1) + 3) // 4)).sum())
    rows = int((cnt.max() + 3) // 4)
    if rows > 8 or 64 * rows * ns > max_growth * max(per_slice, 1):
        return 0
    return max(rows, 1)


def packed_rows(rows: np.ndarray, n_verts: int):
    """Level-0 row vertices packed one 64-bit word per row for the factored tet passes
    (mgpbd_mg_set_level0_packed_rows): the first vertex in the low bb bits, the other three as
    signed ob-bit differences from it. Returns (words uint64, bb, ob), or None when a
    difference does not fit (the passes then read the int4 rows)."""
    r = np.asarray(rows, np.int64).reshape(-1, 4)
    bb = max(1, int(max(n_verts, 1) - 1).bit_length())
    ob = min(20, (64 - bb) // 3)
    if bb > 31 or ob < 2:
        return None
    d = r[:, 1:] - r[:, :1]
    lim = 1 << (ob - 1)
    if len(r) and (d.min() < -lim or d.max() >= lim):
        return None
    mask = (1 << ob) - 1
    w = r[:, 0].astype(np.uint64)
    for k in range(3):
        w |= (d[:, k] & mask).astype(np.uint64) << np.uint64(bb + k * ob)
    return w, bb, ob


def vertex_slices(vptr: np.ndarray, vinc: np.ndarray, rows_uniform: int = 0):
    """Vertex-sliced incidence of the factored level 0, four lanes per vertex: vertices binned
    16 per 64-lane slice, slot k of vertex 16 s + l at vs_ptr[s] + 64 (k // 4) + 4 l + k % 4
    for k < vs_cnt[v] (the vertex's incidences 4 tet + local index, in vinc order); a slice
    holds ceil(busiest vertex / 4) rows of 64 slots, or rows_uniform rows each (then
    vs_ptr[s] = 64 rows_uniform s, and the vertex passes need no slice pointer). Returns
    (vs_ptr int32[n_slices + 1], vs_cnt int32[n_verts], vs_inc int32[vs_ptr[-1]], zero where
    unused)."""
    vptr = np.asarray(vptr, np.int64)
    nv = len(vptr) - 1
    cnt = np.diff(vptr)
    ns = (nv + 15) // 16
    padded = np.zeros(16 * ns, np.int64)
    padded[:nv] = cnt
    rows = (padded.reshape(ns, 16).max(axis=1) + 3) // 4 if ns else np.zeros(0, np.int64)
    if rows_uniform:
        if ns and rows.max() > rows_uniform:
            raise ValueError("a vertex has more slots than the uniform layout holds")
        rows = np.full(ns, int(rows_uniform), np.int64)
    vs_ptr = np.zeros(ns + 1, np.int64)
    vs_ptr[1:] = np.cumsum(64 * rows)
    if vs_ptr[-1] >= 2 ** 31:
        raise ValueError("vertex incidence exceeds 32-bit positions")
    v_of = np.repeat(np.arange(nv, dtype=np.int64), cnt)
    k_of = np.arange(len(v_of), dtype=np.int64) - np.repeat(vptr[:-1], cnt)
    pos = vs_ptr[v_of // 16] + 64 * (k_of // 4) + 4 * (v_of % 16) + k_of % 4
    vs_inc = np.zeros(max(int(vs_ptr[-1]), 1), np.int32)
    vs_inc[pos] = np.asarray(vinc, np.int32)[:len(pos)]
    return vs_ptr.astype(np.int32), cnt.astype(np.int32), vs_inc


def slot_offset_base(vptr: np.ndarray, vinc: np.ndarray):
    """Per-vertex base tet (the smallest tet at the vertex) when every vertex's tets lie within
    65535 of it, else None: the factored level 0 then stores each incidence slot's tet as a
    16-bit offset beside its K and gradient (mgpbd_mg_set_level0_offsets)."""
    vptr = np.asarray(vptr, np.int64)
    cnt = np.diff(vptr)
    nv = len(cnt)
    if nv == 0:
        return None
    tet = np.asarray(vinc, np.int64)[:vptr[-1]] >> 2
    v_of = np.repeat(np.arange(nv), cnt)
    base = np.full(nv, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(base, v_of, tet)
    base[cnt == 0] = 0
    if len(tet) and int((tet - base[v_of]).max()) > 65535:
        return None
    return base.astype(np.int32)


def factored_galerkin_maps(t_i, n_verts: int, agg, c_indptr, c_indices):
    """Maps of level 1 from the gradients (A1 = G W G^T + P^T diag(alpha~) P, symmetric: each
    entry (I, J) with I <= J is summed once and stored at both positions). Returns gptr, gmem,
    gvert: the (aggregate, vertex) pairs in ascending (aggregate, vertex) order with their
    incidences 4 tet + local vertex (ascending) and vertex; tptr, terms: per upper entry the
    pairs (p_Iv, p_Jv) of the vertices v that I and J share, in vertex order; cup, clo: the CSR
    index of the entry and of its transpose; cagg: I for a diagonal entry, else -1. Raises when
    the level-1 pattern lacks a shared-vertex entry or holds an entry with no shared vertex
    (the pattern must be the structural one)."""
    tv = np.asarray(t_i, np.int64).reshape(-1, 4)
    nt = tv.shape[0]
    inc = np.arange(4 * nt, dtype=np.int64)
    agg = np.asarray(agg, np.int64)
    key = agg[inc >> 2] * n_verts + tv.reshape(-1)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(len(ks), bool)
    first[1:] = ks[1:] != ks[:-1]
    starts = np.flatnonzero(first)
    pair_key = ks[starts]
    ng = len(pair_key)
    gptr = np.append(starts, len(ks)).astype(np.int64)
    gmem = inc[order]
    gI, gv = pair_key // n_verts, pair_key % n_verts
    pv = np.argsort(gv, kind="stable")          # pairs by vertex, aggregates ascending within
    m = np.bincount(gv, minlength=n_verts).astype(np.int64)
    vstart = np.concatenate([[0], np.cumsum(m)])
    m_each = m[gv[pv]]
    T = int(m_each.sum())
    a_pos = np.repeat(np.arange(ng, dtype=np.int64), m_each)
    offs = np.arange(T, dtype=np.int64) - np.repeat(np.cumsum(m_each) - m_each, m_each)
    b_pos = np.repeat(vstart[gv[pv]], m_each) + offs
    a, b = pv[a_pos], pv[b_pos]
    n1 = len(c_indptr) - 1
    rows = np.repeat(np.arange(n1, dtype=np.int64), np.diff(c_indptr))
    cols = np.asarray(c_indices, np.int64)
    ckey = rows * n1 + cols

    def find(k):
        c = np.searchsorted(ckey, k)
        if np.any(c >= len(ckey)) or np.any(ckey[np.minimum(c, len(ckey) - 1)] != k):
            raise ValueError("level-1 pattern lacks an entry of two aggregates sharing a vertex")
        return c

    cnt_all = np.bincount(find(gI[a] * n1 + gI[b]), minlength=len(ckey))
    if np.any(cnt_all == 0):
        raise ValueError("a level-1 entry has no shared vertex")
    up = gI[a] <= gI[b]
    a, b = a[up], b[up]
    c = find(gI[a] * n1 + gI[b])
    cup = np.flatnonzero(rows <= cols)          # upper entries, CSR order
    slot = np.full(len(ckey), -1, np.int64)
    slot[cup] = np.arange(len(cup))
    e = slot[c]
    o = np.argsort(e, kind="stable")
    cnt = np.bincount(e, minlength=len(cup))
    terms = np.stack([a[o], b[o]], axis=1)
    tptr = np.concatenate([[0], np.cumsum(cnt)])
    if tptr[-1] >= 2 ** 31 or len(gmem) >= 2 ** 31:
        raise ValueError("level-1 maps exceed 32-bit offsets")
    clo = find(cols[cup] * n1 + rows[cup])
    cagg = np.where(rows[cup] == cols[cup], rows[cup], -1)
    return gptr, gmem, gv, tptr, terms, cup, clo, cagg


def _aggs_from_Ps(Ps):
    """Piecewise-constant P (one positive entry per row, equal within each column: the unit P
    of tet aggregation or pyamg's tentative 1/sqrt(|aggregate|) one) -> aggregate maps. A
    column scaling of P does not change the V-cycle, so both run on the unit-P kernels."""
    out = []
    for P in Ps:
        P = sp.csr_matrix(P)
        ok = np.all(np.diff(P.indptr) == 1) and np.all(P.data > 0)
        if ok:
            col = P.indices
            first = np.full(P.shape[1], np.nan)
            first[col[::-1]] = P.data[::-1]
            ok = np.all(P.data == first[col])
        if not ok:
            raise ValueError("GPUMultigridSolver supports piecewise-constant P (one positive "
                             "entry per row, constant within each column)")
        out.append((P.indices.astype(np.int32), P.shape[1]))
    return out


class GPUMultigridSolver:
    """Multigrid-preconditioned CG solver, fully GPU-resident (cons/mgpcg_gpu.py:18)."""

    def __init__(self, n_smooth: int = 2, omega_jacobi: float = 2.0 / 3.0,
                 maxiter_cg: int = 100, tol_cg: float = 1e-5, *, relative: bool = False,
                 coarse_sweeps: int = 50, galerkin: str = "values",
                 coarse_solver: str = "jacobi", precision: str = "fp32",
                 flexible: bool | None = None, cycle: str = "V", overcorrect: float = 1.0,
                 smoother: str = "jacobi", smoother_target: float = 4.0 / 3.0,
                 smoother_safety: float = 1.1, power_iters: int = 2, cheb_ratio: float = 12.0,
                 verify: int = 0, fp64_iterate: bool = False, stall_rescue: bool = True,
                 smooth_schedule=None, omega_levels=None, start_criterion: str = "residual",
                 device=None):
        self.n_smooth = int(n_smooth)
        # fixed per-level Jacobi weights (the last entry repeating), with smoother='jacobi'
        self.omega_levels = None if omega_levels is None else [float(w) for w in omega_levels]
        self.omega_jacobi = float(omega_jacobi)
        self.maxiter_cg = int(maxiter_cg)
        self.tol_cg = float(tol_cg)
        self.relative = bool(relative)
        self.coarse_sweeps = int(coarse_sweeps)
        if galerkin not in ("values", "structural"):
            raise ValueError("galerkin must be 'values' (reference) or 'structural'")
        self.galerkin = galerkin
        if coarse_solver not in ("jacobi", "cholesky"):
            raise ValueError("coarse_solver must be 'jacobi' (reference) or 'cholesky'")
        self.coarse_solver = coarse_solver
        if precision not in ("fp32", "fp16", "bf16"):
            raise ValueError("precision must be 'fp32', 'fp16' or 'bf16'")
        self.precision = precision
        self.flexible = (precision != "fp32") if flexible is None else bool(flexible)
        if cycle not in ("V", "W"):
            raise ValueError("cycle must be 'V' (reference) or 'W'")
        self.cycle = cycle
        self.overcorrect = float(overcorrect)
        if smoother not in ("jacobi", "spectral", "chebyshev"):
            raise ValueError("smoother must be 'jacobi' (reference), 'spectral' or 'chebyshev'")
        self.smoother = smoother
        if self.omega_levels is not None and (smoother != "jacobi" or not self.omega_levels):
            raise ValueError("omega_levels (a non-empty list) goes with smoother='jacobi'")
        self.smoother_target = float(smoother_target)
        self.smoother_safety = float(smoother_safety)
        self.power_iters = int(power_iters)
        self.cheb_ratio = float(cheb_ratio)
        self.verify = int(verify)   # 0 off, 1 verify + restart (True), 2 measure only
        self.fp64_iterate = bool(fp64_iterate)
        self.stall_rescue = bool(stall_rescue)
        if start_criterion not in ("residual", "energy"):
            raise ValueError("start_criterion must be 'residual' or 'energy'")
        self.start_criterion = start_criterion
        # smoothing steps per level (the last entry repeats); None: n_smooth everywhere
        self.smooth_schedule = None if smooth_schedule is None else [int(k) for k in smooth_schedule]
        self.stall_ratio = 0.25   # squared-residual ratio over 8 iterations counted as a stall
        self.last_lambda = None
        self.n_levels = 0
        self._initialized = False
        # the hierarchy, PCG state and PCG graph live in the context: one per solver, so that
        # several solvers on one device do not overwrite each other's levels
        self.ctx = new_context(device)
        self.device = self.ctx.device
        self.last_info = None
        self._fac_src = None
        self._fac = None

    def set_factored_level0(self, t_i, v_invm, gradC, alpha_tilde, vptr, vinc, t_i_host=None):
        """Level 0 in factored form (extension; see the module notes). Call before ``setup``.
        The device tensors t_i (int32, 4 per tet), v_invm, gradC (12 floats per tet) and
        alpha_tilde are read at every ``update_all_coarse_matrices`` and every solve and must
        stay allocated; vptr / vinc: the host vertex incidence (4 tet + local index)."""
        if self.precision == "fp32":
            raise ValueError("the factored level 0 is a reduced-precision option")
        if self.galerkin != "structural":
            raise ValueError("the factored level 0 needs galerkin='structural' (level 1 from "
                             "the gradients fills every shared-vertex entry)")
        if self._initialized:
            raise RuntimeError("set_factored_level0 must precede setup()")
        th = t_i.cpu().numpy() if t_i_host is None else np.asarray(t_i_host)
        self._fac_src = dict(t_i=t_i, v_invm=v_invm, gradC=gradC, alpha_tilde=alpha_tilde,
                             vptr=np.asarray(vptr), vinc=np.asarray(vinc),
                             t_i_host=np.ascontiguousarray(th, np.int32).reshape(-1))

    @property
    def factored(self) -> bool:
        """Level 0 runs factored (its stored values are not read by the solve)."""
        return self._fac is not None

    def setup(self, A_level0, Ps_or_aggs: list):
        """Hierarchy from the fine matrix and prolongations (cons/mgpcg_gpu.py:30-173).

        ``A_level0``: scipy CSR (values used for the pattern as in the reference) or a
        ``GPUSparseMatrix`` already holding level-0 values (used in place, no copy).
        ``Ps_or_aggs``: scipy P matrices or ``(agg, n_coarse)`` pairs."""
        lib = _lib.load()
        if isinstance(A_level0, GPUSparseMatrix):
            A0 = A_level0
        else:
            A = sp.csr_matrix(A_level0)
            A0 = GPUSparseMatrix(A.shape[0], A.nnz, device=self.device.index)
            A0.from_scipy(A)
        if Ps_or_aggs and not isinstance(Ps_or_aggs[0], tuple):
            aggs = _aggs_from_Ps(Ps_or_aggs)
        else:
            aggs = [(np.ascontiguousarray(a, np.int32), int(n)) for a, n in Ps_or_aggs]
        self.levels = [A0]
        self.transfers = []
        self.refill = []
        fine_vals = A0.csr_data() if self.galerkin == "values" else None
        F = A0
        for agg, n_c in aggs:
            assert agg.shape[0] == F.n_rows
            vals = None if fine_vals is None else np.ascontiguousarray(fine_vals, np.float32)
            g = lib.mgpbd_galerkin_build(F.n_rows, np_ptr(F.indptr), np_ptr(F.indices),
                                         None if vals is None else vals.ctypes.data_as(_lib.F32P),
                                         np_ptr(agg), n_c)
            if not g:
                check(-1, "mgpbd_galerkin_build")
            cnnz, ncon = ctypes.c_int64(), ctypes.c_int64()
            check(lib.mgpbd_galerkin_sizes(g, ctypes.byref(cnnz), ctypes.byref(ncon)),
                  "mgpbd_galerkin_sizes")
            c_indptr = np.empty(n_c + 1, np.int32)
            c_indices = np.empty(max(cnnz.value, 1), np.int32)
            c_diag = np.empty(n_c, np.int32)
            cidx = np.empty(max(ncon.value, 1), np.int32)
            coff = np.empty(cnnz.value + 1, np.int32)
            check(lib.mgpbd_galerkin_copy(g, np_ptr(c_indptr), np_ptr(c_indices),
                                          np_ptr(c_diag), np_ptr(cidx), np_ptr(coff)),
                  "mgpbd_galerkin_copy")
            lib.mgpbd_galerkin_free(g)
            c_indices, cidx = c_indices[:cnnz.value], cidx[:ncon.value]
            C = GPUSparseMatrix(n_c, int(cnnz.value), device=self.device.index)
            C.set_pattern(c_indptr, c_indices, diag_csr_pos=c_diag)
            dev = self.device
            fpos = F.sell_pos_host[cidx] if len(cidx) else np.zeros(1, np.int32)
            self.refill.append({
                "n": int(cnnz.value),
                "off": torch.from_numpy(coff).to(dev),
                "fpos": torch.from_numpy(np.ascontiguousarray(fpos, np.int32)).to(dev),
                "cpos": torch.from_numpy(np.ascontiguousarray(C.sell_pos_host, np.int32)).to(dev),
            })
            mptr, mem = _members(agg, n_c)
            self.transfers.append({
                "agg": torch.from_numpy(agg).to(dev),
                "mptr": torch.from_numpy(mptr).to(dev),
                "mem": torch.from_numpy(mem).to(dev),
                "contrib_idx": cidx, "contrib_off": coff,
            })
            self.levels.append(C)
            if fine_vals is not None:   # coarse values for the next level's pattern
                fine_vals = self._host_refill(fine_vals, cidx, coff)
            F = C
        self.n_levels = len(self.levels)
        self._level_sizes = [L.n_rows for L in self.levels]
        # work vectors per level: b, x, x', r, two W-cycle and two power-iteration buffers
        self._work = [torch.zeros(8 * 64 * ((L.n_rows + 63) // 64), dtype=torch.float32,
                                  device=self.device) for L in self.levels]
        n0 = self._level_sizes[0]
        # r, z, p, Ap and the FP64 iterate
        self._pcg_work = torch.zeros(6 * n0, dtype=torch.float32, device=self.device)
        self._b = [w[:L.n_rows] for w, L in zip(self._work, self.levels)]
        self._x = [torch.zeros(L.n_rows, dtype=torch.float32, device=self.device)
                   for L in self.levels]
        self._configure()
        self._setup_half()
        self.update_all_coarse_matrices()
        self._initialized = True

    # reduced-precision levels pack 16-bit column offsets j - i with the values when every
    # |j - i| is below this bound, and use int32 columns (shared with the FP32 copy) otherwise
    col16_limit = 32768
    # factored level 0: 16-bit tet offsets in the vertex slots' unused words when they fit, and
    # then (when it costs at most 25 % more slots) every vertex slice as tall as the busiest
    # vertex needs
    slot_offsets = True
    # factored level 0: every row's four vertices in one 64-bit word for the tet passes when
    # the differences to the first vertex fit
    pack_rows = True
    # packed levels whose blocks of 256 rows read x from at most this many 32-column segments
    # stage them in LDS (windowed sweeps; the words then hold LDS offsets, so the 16-bit column
    # offset bound does not apply); 0 disables
    window_segs_limit = 384

    def _window(self, L):
        """(seg_ptr, segs, largest window) of a level's LDS windows, None above the limit."""
        if self.window_segs_limit <= 0 or L.n_rows == 0:
            return None
        lib = _lib.load()
        nb = (64 * ((L.n_rows + 63) // 64) + 255) // 256
        seg_ptr = np.zeros(nb + 1, np.int32)
        check(lib.mgpbd_sell_window(L.n_rows, np_ptr(L.indptr), np_ptr(L.indices),
                                    np_ptr(seg_ptr), None, 0), "mgpbd_sell_window")
        widest = int(np.diff(seg_ptr).max())
        if widest > min(self.window_segs_limit, 384):
            return None
        segs = np.zeros(max(int(seg_ptr[-1]), 1), np.int32)
        check(lib.mgpbd_sell_window(L.n_rows, np_ptr(L.indptr), np_ptr(L.indices),
                                    np_ptr(seg_ptr), np_ptr(segs), len(segs)), "mgpbd_sell_window")
        return seg_ptr, segs, widest

    def _setup_half(self):
        """Reduced-precision copies of every level for the V-cycle (extension), and the LDS
        windows: of every packed level, and of level 0 for the PCG's FP32 kernels."""
        prec = {"fp32": 0, "fp16": 1, "bf16": 2}[self.precision]
        self.ctx("mgpbd_mg_set_precision", prec, int(self.flexible))
        self._half = []
        dev = self.device
        wins = [self._window(L) for L in self.levels]
        self._win = []
        if prec == 0:
            for l, win in enumerate(wins):
                if win is not None:
                    self._set_window(l, win, hpk=False)
            return
        for l, L in enumerate(self.levels):
            if l == 0 and self._fac_src is not None:
                self._setup_factored(wins[0])
                continue
            rows = np.repeat(np.arange(L.n_rows, dtype=np.int64), np.diff(L.indptr))
            delta = L.indices.astype(np.int64) - rows
            win = wins[l]
            use16 = bool(win is not None or len(delta) == 0 or
                         np.abs(delta).max() < self.col16_limit)
            slots = max(L.padded, 1)
            hb = {
                # packed (col - row) << 16 | half words, or half values beside int32 columns
                "hpk": torch.zeros(slots, dtype=torch.int32, device=dev) if use16 else None,
                "hval": None if use16 else torch.zeros(slots, dtype=torch.int16, device=dev),
                "sd": torch.ones(L.n_rows, dtype=torch.float32, device=dev),
                "isd": torch.ones(L.n_rows, dtype=torch.float32, device=dev),
            }
            self.ctx("mgpbd_mg_set_level_half", l, dptr(hb["hval"]), dptr(hb["hpk"]),
                     dptr(hb["sd"]), dptr(hb["isd"]), dptr(L.diag_pos))
            hb["seg_ptr"] = hb["segs"] = None
            self._half.append(hb)
            if win is not None:
                self._set_window(l, win, hpk=True)

    def _setup_factored(self, win):
        """Level 0 without stored half values or LDS window (nothing reads its stored values
        while it runs factored): scale arrays, the factored operators' buffers and level 1's
        maps from the gradients."""
        dev = self.device
        L = self.levels[0]
        src = self._fac_src
        hb = {"hpk": None, "hval": None,
              "sd": torch.ones(L.n_rows, dtype=torch.float32, device=dev),
              "isd": torch.ones(L.n_rows, dtype=torch.float32, device=dev),
              "seg_ptr": None, "segs": None}
        self.ctx("mgpbd_mg_set_level_half", 0, None, None, dptr(hb["sd"]), dptr(hb["isd"]),
                 dptr(L.diag_pos))
        self._half.append(hb)
        tbase = slot_offset_base(src["vptr"], src["vinc"]) if self.slot_offsets else None
        rows_u = uniform_rows(src["vptr"]) if tbase is not None else 0
        vs_ptr, vs_cnt, vs_inc = vertex_slices(src["vptr"], src["vinc"], rows_u)
        nv, nt = len(vs_cnt), L.n_rows
        f = self._fac = {
            "vs_ptr": torch.from_numpy(vs_ptr).to(dev),
            "vs_cnt": torch.from_numpy(vs_cnt).to(dev),
            "vs_inc": torch.from_numpy(vs_inc).to(dev),
            "vs_k": torch.zeros(2 * len(vs_inc), dtype=torch.int32, device=dev),
            "vs_g": torch.zeros(4 * len(vs_inc), dtype=torch.float32, device=dev),
            "tk": torch.zeros((9 * nt + 1) // 2, dtype=torch.int32, device=dev),
            "ahat": torch.zeros(nt, dtype=torch.float32, device=dev),
            "u": torch.zeros(4 * nv, dtype=torch.float32, device=dev),
            "u64": torch.zeros(4 * nv, dtype=torch.float64, device=dev),
        }
        self.ctx("mgpbd_mg_set_level0_factored", nv, dptr(src["t_i"]), dptr(src["v_invm"]),
                 dptr(src["gradC"]), dptr(src["alpha_tilde"]), dptr(f["vs_ptr"]),
                 dptr(f["vs_cnt"]), dptr(f["vs_inc"]), dptr(f["vs_k"]), dptr(f["vs_g"]),
                 dptr(f["tk"]), dptr(f["ahat"]), dptr(f["u"]), dptr(f["u64"]),
                 dptr(L.diag_inv))
        if tbase is not None:
            f["tbase"] = torch.from_numpy(tbase).to(dev)
            f["rows_uniform"] = rows_u
            self.ctx("mgpbd_mg_set_level0_offsets", dptr(f["tbase"]), rows_u)
        pk = packed_rows(src["t_i_host"], nv) if self.pack_rows else None
        if pk is not None:
            f["rows_packed"] = torch.from_numpy(pk[0].view(np.int64)).to(dev)
            f["rows_bits"] = pk[1:]
            self.ctx("mgpbd_mg_set_level0_packed_rows", dptr(f["rows_packed"]), pk[1], pk[2])
        if self.n_levels < 2:
            return
        C = self.levels[1]
        gptr, gmem, gvert, tptr, terms, cup, clo, cagg = factored_galerkin_maps(
            src["t_i_host"], nv, self.transfers[0]["agg"].cpu().numpy(), C.indptr, C.indices)
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        f.update(gptr=up(gptr), gmem=up(gmem), gvert=up(gvert), tptr=up(tptr),
                 terms=up(terms.reshape(-1)), cagg=up(cagg), cpos=up(C.sell_pos_host[cup]),
                 cpos_t=up(C.sell_pos_host[clo]),
                 G=torch.zeros(4 * max(len(gvert), 1), dtype=torch.float32, device=dev))
        self.ctx("mgpbd_mg_set_level1_factored", len(gvert), dptr(f["gptr"]), dptr(f["gmem"]),
                 dptr(f["gvert"]), dptr(f["G"]), len(cagg), dptr(f["tptr"]), dptr(f["terms"]),
                 dptr(f["cpos"]), dptr(f["cpos_t"]), dptr(f["cagg"]), dptr(C.data))

    def level_residual(self, level: int, b: torch.Tensor, x: torch.Tensor, r: torch.Tensor):
        """r = b - A_l x with the operator the V-cycle applies on level l (scaled variables in
        reduced precision)."""
        self.ctx("mgpbd_mg_level_residual", int(level), dptr(b), dptr(x), dptr(r))

    def _set_window(self, l, win, hpk):
        """Uploads level l's window lists and switches its kernels to the windowed ones. The
        16-bit offsets of the FP32 kernels (woff) go to level 0, whose FP32 matrix the PCG
        multiplies, and to every level of an FP32 V-cycle."""
        dev = self.device
        L = self.levels[l]
        seg_ptr = torch.from_numpy(win[0]).to(dev)
        segs = torch.from_numpy(win[1]).to(dev)
        fp32 = l == 0 or not hpk
        woff = torch.zeros(max(L.padded, 1), dtype=torch.int16, device=dev) if fp32 else None
        self.ctx("mgpbd_mg_set_level_window", l, dptr(seg_ptr), dptr(segs), len(win[0]) - 1,
                 win[2], dptr(woff))
        if hpk:
            self._half[l]["seg_ptr"], self._half[l]["segs"] = seg_ptr, segs
        self._win.append({"level": l, "seg_ptr": seg_ptr, "segs": segs, "woff": woff})

    @staticmethod
    def _host_refill(fine_vals, cidx, coff):
        """Coarse values on the host (FP64 sums of the FP32 fine values, in map order)."""
        nc = len(coff) - 1
        seg = np.repeat(np.arange(nc), np.diff(coff))
        out = np.bincount(seg, weights=fine_vals[cidx].astype(np.float64), minlength=nc)
        return out.astype(np.float32)

    def _configure(self):
        ctx = self.ctx
        mode = 1 if (self.coarse_solver == "cholesky" and self.n_levels > 1) else 0
        ctx("mgpbd_mg_configure", self.n_levels, self.n_smooth, self.omega_jacobi, mode,
            self.coarse_sweeps)
        ctx("mgpbd_mg_set_cycle", 2 if self.cycle == "W" else 1, self.overcorrect)
        ctx("mgpbd_mg_set_smoother", {"jacobi": 0, "spectral": 1, "chebyshev": 2}[self.smoother],
            self.smoother_target, self.smoother_safety, self.power_iters, self.cheb_ratio)
        if self.omega_levels:
            w = np.asarray(self.omega_levels, np.float32)
            ctx("mgpbd_mg_set_level_weights", len(w), w.ctypes.data_as(_lib.F32P))
        ctx("mgpbd_pcg_set_verify", int(self.verify))
        ctx("mgpbd_pcg_set_fp64_iterate", int(self.fp64_iterate))
        ctx("mgpbd_pcg_set_start_criterion", int(self.start_criterion == "energy"))
        ctx("mgpbd_pcg_set_rescue", int(self.stall_rescue and self.smoother != "jacobi"), 1.25,
            self.stall_ratio)
        if self.smooth_schedule:
            for l in range(self.n_levels):
                k = self.smooth_schedule[min(l, len(self.smooth_schedule) - 1)]
                ctx("mgpbd_mg_set_level_smoothing", l, k)
        self._coarse_mode = mode
        for l, rf in enumerate(self.refill):   # for mgpbd_mg_refresh (the native substep)
            F, C = self.levels[l], self.levels[l + 1]
            ctx("mgpbd_mg_set_refill", l, rf["n"], dptr(rf["off"]), dptr(rf["fpos"]),
                dptr(rf["cpos"]), dptr(F.data), dptr(C.data), C.n_rows, dptr(C.diag_pos),
                dptr(C.diag_inv))
        for l, L in enumerate(self.levels):
            t = self.transfers[l] if l < len(self.transfers) else None
            ctx("mgpbd_mg_set_level", l, L.n_rows, dptr(L.slice_ptr), dptr(L.sell_idx),
                dptr(L.data), dptr(L.diag_inv),
                dptr(t["agg"]) if t else None, dptr(t["mptr"]) if t else None,
                dptr(t["mem"]) if t else None, dptr(self._work[l]), int(L.is_diagonal))

    # -- per-iteration updates -------------------------------------------------------------
    def update_fine_matrix_data(self, data):
        """New level-0 values (cons/mgpcg_gpu.py:175-181): numpy in CSR order, or a device
        tensor in CSR order; refreshes 1/diag."""
        self.levels[0].update_data(data)
        self.levels[0].cache_diag_inv()

    def update_all_coarse_matrices(self, refresh_smoother: bool = True, after_level0=None,
                                   coarse: bool = True):
        """Coarse values from the fine ones through the refill maps
        (cons/mgpcg_gpu.py:183-212). refresh_smoother=False keeps the spectral smoother's
        current estimates and the coarsest level's dense inverse (extension: the caller knows
        the matrices changed little; the kept inverse is SPD, so the V-cycle stays an SPD
        preconditioner). coarse=False keeps levels 1.. as they are and refreshes only level
        0's reduced-precision data (extension, mgpbd_mg_refresh with coarse = 0).
        after_level0 (factored level 0 only): called once level 0's scales, 1/diag and K are
        queued, before the coarse levels' work, so the caller can start work that needs only
        those on another stream."""
        early = after_level0 is not None and self._fac is not None
        if early:
            self.ctx("mgpbd_mg_convert_half_levels", 0, 1)
            after_level0()
        if not coarse:
            if getattr(self, "_half", None) and not early:
                self.ctx("mgpbd_mg_convert_half_levels", 0, 1)
            self.ctx("mgpbd_mg_set_coarse_lagged", 1)
            if self.smoother != "jacobi" and refresh_smoother:
                self.ctx("mgpbd_mg_update_smoother", None)
            return
        for l, rf in enumerate(self.refill):
            F, C = self.levels[l], self.levels[l + 1]
            if l == 0 and self._fac is not None:
                self.ctx("mgpbd_mg_factored_refill")
            elif rf["n"] > 0:
                self.ctx("mgpbd_coarse_refill", rf["n"], dptr(rf["off"]), dptr(rf["fpos"]),
                         dptr(rf["cpos"]), dptr(F.data), dptr(C.data))
            C.cache_diag_inv()
        if getattr(self, "_coarse_mode", 0) == 1 and refresh_smoother:
            self.ctx("mgpbd_mg_factor_coarsest")
        if getattr(self, "_half", None):
            if early:
                self.ctx("mgpbd_mg_convert_half_levels", 1, self.n_levels)
            else:
                self.ctx("mgpbd_mg_convert_half")
        self.ctx("mgpbd_mg_set_coarse_lagged", 0)
        if self.smoother != "jacobi" and refresh_smoother:
            self.ctx("mgpbd_mg_update_smoother", None)

    def device_bytes(self, seen=None) -> int:
        """HBM held by the hierarchy: level matrices (FP32 and reduced-precision copies),
        transfers, refill maps, work vectors and the context's own buffers."""
        from .runtime import device_bytes
        ctx_bytes = ctypes.c_int64()
        self.ctx("mgpbd_ctx_device_bytes", ctypes.byref(ctx_bytes))
        return device_bytes(self.levels, self.transfers, self.refill, self._work,
                            self._pcg_work, self._x, getattr(self, "_half", None),
                            getattr(self, "_win", None), self._fac, seen=seen) + ctx_bytes.value

    def pcg_counters(self):
        """(rescues: stall and lagged-hierarchy, verification restarts) over this solver's
        solves."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self.ctx("mgpbd_pcg_counters", ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def lag_rescues(self) -> int:
        """Solves that rebuilt a lagged hierarchy mid-solve (mgpbd_mg_set_coarse_lagged)."""
        a = ctypes.c_int64()
        self.ctx("mgpbd_pcg_lag_rescues", ctypes.byref(a))
        return int(a.value)

    def coarse_stats(self):
        """(Newton-Schulz refresh steps kept, full inversions) of the coarsest inverse."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self.ctx("mgpbd_mg_coarse_stats", ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def smoother_weights(self, chebyshev: bool = False):
        """Jacobi weight of each level as the V-cycle uses it now (synchronises); with
        ``chebyshev`` also the (c1, c2) of Chebyshev steps 2..17, shape (n_levels, 16, 2)."""
        w = np.zeros(self.n_levels, np.float32)
        c = np.zeros((self.n_levels, 16, 2), np.float32)
        self.ctx("mgpbd_mg_get_smoother", w.ctypes.data_as(_lib.F32P),
                 c.ctypes.data_as(_lib.F32P) if chebyshev else None)
        return (w, c) if chebyshev else w

    def smoother_estimates(self):
        """Largest eigenvalue of D^-1 A per level from the spectral smoother's power iteration
        (advances it by its usual iteration count; synchronises)."""
        lam = np.zeros(self.n_levels, np.float64)
        self.ctx("mgpbd_mg_update_smoother", lam.ctypes.data_as(_lib.F64P))
        self.last_lambda = lam
        return lam

    # -- solve -----------------------------------------------------------------------------
    def vcycle(self, level: int = 0):
        """self._x[0] = V-cycle(self._b[0]) from zero (cons/mgpcg_gpu.py:214-259)."""
        if level != 0:
            raise ValueError("only the level-0 V-cycle is exposed")
        self.ctx("mgpbd_mg_vcycle", dptr(self._b[0]), dptr(self._x[0]))

    def apply(self, r: torch.Tensor, z: torch.Tensor):
        """z = M^-1 r (one V-cycle)."""
        self.ctx("mgpbd_mg_vcycle", dptr(r), dptr(z))

    def solve(self, b_field: torch.Tensor, x_field: torch.Tensor) -> int:
        """PCG for A x = b from x = 0; returns cg_iters (cons/mgpcg_gpu.py:261-329)."""
        it = ctypes.c_int(0)
        info = np.zeros(4, np.float64)
        self.ctx("mgpbd_pcg_solve", dptr(b_field), dptr(x_field), dptr(self._pcg_work),
                 self.maxiter_cg, self.tol_cg, int(self.relative), ctypes.byref(it),
                 info.ctypes.data_as(_lib.F64P))
        self.last_info = {"rr": float(info[0]), "bb": float(info[1]), "reason": int(info[2]),
                          "rr_true": float(info[3])}
        return int(it.value)

    def solve_guess(self, b_field: torch.Tensor, x_field: torch.Tensor,
                    x_alt: torch.Tensor | None = None) -> int:
        """PCG from the guess held in ``x_field`` (extension): the device takes whichever of
        the guess, ``x_alt`` and zero has the smallest residual ||b - A x0|| (or, with
        start_criterion='energy', the lowest CG energy x0.A x0 / 2 - b.x0), then iterates to
        the same stopping test as ``solve``. ``last_info['start']`` is 0 (zero), 1 (guess) or
        2 (x_alt)."""
        it = ctypes.c_int(0)
        info = np.zeros(8, np.float64)
        self.ctx("mgpbd_pcg_solve_guess", dptr(b_field), dptr(x_field), dptr(x_alt),
                 dptr(self._pcg_work), self.maxiter_cg, self.tol_cg, int(self.relative),
                 ctypes.byref(it), info.ctypes.data_as(_lib.F64P))
        self.last_info = {"rr": float(info[0]), "bb": float(info[1]), "reason": int(info[2]),
                          "rr0": float(info[3]), "start": int(info[4]),
                          "rr_guess": float(info[5]), "rr_alt": float(info[6]),
                          "rr_true": float(info[7])}
        return int(it.value)
