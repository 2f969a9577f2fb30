"""Benchmark: MGPBD substeps/s on one MI355X.

Default workload (the 1M-tet configuration, controllable_pbd_3d_amd/workloads.py "c1"):
jittered block of 56^3 Kuhn cubes (185,193 vertices, 1,053,696 tets), h = 0.02 m, per-tet E
log-uniform in [1e4, 1e6] Pa (seed 0), nu = 0.45, rho = 1000 kg/m^3, gravity -9.8 m/s^2, top
face held, dt = 1/300 s (60 fps x 5 substeps), Newton defaults of the reference (maxiter 20,
atol 1e-4, rtol 1e-2, line search), PCG to relative residual 1e-6 (max 100 iterations),
tet-aggregation hierarchy extended to 5 levels, coarsest level solved exactly (FP64 dense
inverse). One ARAP constraint per tet (the reference's material) unless a run says "volume".

Runs (each its own solver on the same mesh, same loading):
  fp32_reference  the reference's algorithm (2 + 2 weighted-Jacobi sweeps, omega 2/3, FP32,
                  assembled CSR; the FP64 residual is measured, not enforced)
  fp32            Chebyshev smoothing on the device's spectral estimates, FP64 iterate, every
                  solve FP64-verified; FP32 assembled level 0 (the metric's denominator);
                  spectral estimates and coarsest inverse rebuilt at each substep's first
                  Newton iteration, coarse levels 1.. then and after every solve of >= 8 PCG
                  iterations (level 0 at every one), as in every run below
  fp16_assembled  the same with the stored FP16 level 0 (same operator as fp32: the precision's
                  own gain)
  fp16, bf16      FP16 / BF16 V-cycle on the factored level 0 J W J^T + alpha~
  fp16_previous   fp16 started from the previous substep's lambda
  fp16_learned    fp16 started from the learned warm start (network trained in the setup)
  fp32_volume, fp16_volume   fp32 / fp16 with the volume constraint beside ARAP (lambda from
                  nu = 0.45: 2 constraints per tet)
A step is one substep (predict, Newton/MGPCG solve, velocity update) from the state the
previous step left; inputs are resident in HBM. Substeps run in the library (mgpbd_substep, the
C++ Newton loop); `--python-loop` runs the Python Newton loop instead (same kernels, same
results bit for bit).

Prints one JSON line. `--gpus N` with N > 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from controllable_pbd_3d_amd.workloads import (CONFIGS, Loading, build_workload,   # noqa: E402
                                               solver_args)

HBM_PEAK_GBS = 8000.0
NATIVE_SUBSTEP = True      # --python-loop sets False


def kernel_bytes(n_rows, nnz, kind, per_nnz=8, chebyshev=False):
    """Algorithmic HBM bytes per launch of a stored level; per_nnz = 8 for FP32 values + int32
    columns, 6 for FP32 values + 16-bit LDS offsets (windowed level 0), 4 for FP16/BF16 values +
    16-bit column or LDS offsets (each array counted once). With the Chebyshev smoother the
    level-0 sweeps also read the previous iterate (all steps but the Jacobi start): 4 n bytes
    more, counted on every sweep."""
    if kind == "spmv":      # values + columns + row pointers + x + y
        return per_nnz * nnz + 4 * (n_rows + 1) + 4 * n_rows + 4 * n_rows
    if kind == "jacobi":    # + b, D^-1 (FP32 only: the scaled levels have unit diagonal), x -> x_new
        return per_nnz * nnz + 4 * (n_rows + 1) + (16 if per_nnz >= 6 else 12) * n_rows + \
            (4 * n_rows if chebyshev else 0)
    if kind == "resid":     # values + columns + ptr + x + b -> r
        return per_nnz * nnz + 4 * (n_rows + 1) + 12 * n_rows
    raise ValueError(kind)


def factored_bytes(n_rows, n_verts, kind, chebyshev=False, offsets=False, packed=False):
    """Algorithmic HBM bytes of the factored level 0 (each array counted once per pass; 4
    incidence slots per row). kind "gather" (vertex pass): 4 slots per row x (4-byte incidence
    + 8-byte K) + x (4 per row) + the vertex weight read and u written (4 + 12 per vertex) =
    52 n + 16 n_v; with 16-bit slot offsets (the tet in K's fourth word, a 4-byte base tet per
    vertex) 36 n + 20 n_v. "apply" (tet pass of a sweep): t_i 16 + K 18 (three vertices' nine
    halves; the fourth vertex's is their negative sum) + a^ 4 + x 4 + b 4 + y 4 per row + u
    read (12 per vertex) = 50 n + 12 n_v. "jacobi"/"resid": both passes (+ the previous
    iterate, 4 n, for the Chebyshev sweeps). "spmv": the PCG's FP64 product from the FP32 gradients, 4 slots x
    (4-byte incidence + 16-byte gradient; 16 bytes with offsets, + 4 per vertex) + p + u (32
    per vertex) written; t_i 16 + gradC 48 + alpha~ 4 + p 4 + A p 4 per row + u read. With
    packed rows (packed) the tet passes read 8 bytes of row vertices instead of t_i's 16."""
    gather = (36 * n_rows + 20 * n_verts) if offsets else (52 * n_rows + 16 * n_verts)
    tv = 8 if packed else 16
    apply = (34 + tv) * n_rows + 12 * n_verts
    if kind == "gather":
        return gather
    if kind == "apply":
        return apply
    if kind == "spmv":
        vert = (64 * n_rows + 4 * n_verts) if offsets else 80 * n_rows
        return vert + 4 * n_rows + 32 * n_verts + (60 + tv) * n_rows + 32 * n_verts
    if kind in ("jacobi", "resid"):
        return gather + apply + (4 * n_rows if chebyshev and kind == "jacobi" else 0)
    raise ValueError(kind)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel of this workload from the committed rocprofv3 counter
    summaries (``bench-traffic: <workload> <kernel> <bytes>`` lines in profiles/*.md), or None
    when that workload was not profiled."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            parts = line.split()
            if len(parts) == 4 and parts[0] == "bench-traffic:" and parts[1] == workload and \
                    parts[2] == kernel:
                return float(parts[3])
    return None


def kernel_stats(ctx, n, nnz, precision, chebyshev=False, spmv_per_nnz=8, factored_nv=None,
                 offsets=False, packed=False):
    from controllable_pbd_3d_amd import _lib
    ms = np.zeros(8, np.float64)
    cnt = np.zeros(8, np.int64)
    ctx("mgpbd_ctx_profile_read", ms.ctypes.data_as(_lib.F64P),
        cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    kern = {}
    # the outer PCG SpMV is FP32 (values beside int32 columns, or beside 16-bit LDS offsets on
    # a windowed level 0); the V-cycle passes use the precision under test on the same layout
    vb = spmv_per_nnz if precision == "fp32" else 4
    for kind, slot, per in (("spmv", 0, spmv_per_nnz), ("jacobi", 1, vb), ("resid", 2, vb)):
        if cnt[slot] > 0:
            avg = ms[slot] / cnt[slot] * 1e-3
            if factored_nv:
                by = factored_bytes(n, factored_nv, kind, chebyshev, offsets, packed)
            else:
                by = kernel_bytes(n, nnz, kind, per, chebyshev)
            kern[kind] = {"avg_us": avg * 1e6, "launches": int(cnt[slot]),
                          "total_ms": float(ms[slot]), "GBs": by / avg / 1e9, "bytes": by}
    for kind, slot in (("fill_A", 3), ("constraints", 4), ("refill", 5), ("coarse", 6)):
        if cnt[slot] > 0:
            kern[kind] = {"avg_us": ms[slot] / cnt[slot] * 1e3, "launches": int(cnt[slot]),
                          "total_ms": float(ms[slot])}
    return kern


CHEBYSHEV = dict(smoother="chebyshev", n_smooth=5, smooth_schedule=[5, 1], cheb_ratio=30.0,
                 smoother_safety=1.2, verify_residual=1, fp64_iterate=True,
                 spectral_refresh="substep", coarse_refresh="adaptive")
# Level-0 / coarse Chebyshev steps and interval ratio of the runs whose measured best differs
# from the workload's (at 1M tets: 13 + 13 level-0 steps for FP32 and FP16 with the ARAP rows,
# 8 + 8 for BF16, whose 8-bit mantissa caps what longer polynomials gain, and for the volume
# rows; every run is measured at its own best, the FP32 denominator included)
SHORT_SMOOTHING = ((8, 1), 40.0)
RUNS = {
    "fp32_reference": dict(precision="fp32", smoother="jacobi", verify_residual=2),
    "fp32": dict(precision="fp32", **CHEBYSHEV),
    "fp16_assembled": dict(precision="fp16", level0="assembled", **CHEBYSHEV),
    "fp16": dict(precision="fp16", level0="factored", **CHEBYSHEV),
    "bf16": dict(precision="bf16", level0="factored", smoothing=SHORT_SMOOTHING, **CHEBYSHEV),
    "fp16_previous": dict(precision="fp16", level0="factored", warm_start="previous", **CHEBYSHEV),
    "fp16_learned": dict(precision="fp16", level0="factored", warm_start="learned", **CHEBYSHEV),
    "fp32_volume": dict(precision="fp32", volume=True, smoothing=SHORT_SMOOTHING, **CHEBYSHEV),
    "fp16_volume": dict(precision="fp16", level0="factored", volume=True,
                        smoothing=SHORT_SMOOTHING, **CHEBYSHEV),
}
DEFAULT_RUNS = list(RUNS)


def run_kwargs(name, cfg):
    kw = dict(RUNS[name])
    own = kw.pop("smoothing", None)
    sm = own if own is not None and cfg.get("smoothing") else cfg.get("smoothing")
    if kw.get("smoother") == "chebyshev" and sm:
        (k0, k1), ratio = sm
        kw.update(n_smooth=max(k0, k1), smooth_schedule=[k0, k1], cheb_ratio=ratio)
    if "coarse_refresh" in kw and cfg.get("coarse_refresh"):
        kw["coarse_refresh"] = cfg["coarse_refresh"]
    return kw


def make_solver(mesh, mu_t, lam_t, cfg, kw, model=None, record=None):
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    kw = dict(kw)
    volume = kw.pop("volume", False)
    args = solver_args(mesh, mu_t, lam_t, volume=volume,
                       maxiter_cg=cfg.get("maxiter_cg", 100), hierarchy="extended",
                       galerkin="structural", max_coarse=64, max_levels=cfg["max_levels"],
                       coarse_solver="cholesky", warm_start_model=model, record_samples=record,
                       native_substep=NATIVE_SUBSTEP, verbose=False)
    args.update(kw)
    s = MGPBDSolverGPU(**args)
    s.init()
    return s


def train_learned(cfg, seeds, substeps, epochs):
    """Warm-start network trained on first-Newton-iteration pairs the solver produces on
    other seeds of the same workload (mesh jitter and per-tet E differ), A-norm loss."""
    import torch
    from controllable_pbd_3d_amd.warm_start import LearnedWarmStart
    t0 = time.perf_counter()
    samples, model = [], None
    kw = run_kwargs("fp16_previous", cfg)
    for seed in seeds:
        mesh, mu_t, lam_t = build_workload(cfg, seed)
        if model is None:
            model = LearnedWarmStart.for_mesh(mesh.tets, mesh.n_tets, seed=0, lr=1e-3)
        s = make_solver(mesh, mu_t, lam_t, cfg, kw, model=model, record=samples)
        load = Loading(cfg, mesh, seed, substeps)
        for k in range(substeps):
            load.step(s)
        del s
        torch.cuda.empty_cache()
        print(f"[bench] learned: seed {seed} -> {len(samples)} samples", file=sys.stderr,
              flush=True)
    gen_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    before = model.evaluate(samples)
    hist = model.train(samples, epochs=epochs, loss="energy", seed=0)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    info = {"train_seeds": list(seeds), "substeps_per_seed": substeps, "samples": len(samples),
            "epochs": epochs, "energy_loss_untrained": before, "energy_loss_epochs": hist,
            "generate_s": gen_s, "train_s": train_s}
    del samples
    model.release_training_memory()
    torch.cuda.empty_cache()
    return model, info


def probe_level0(s, reps=60, flush_gib=1.0):
    """HIP-event timing of this solver's level-0 sweep kernels (mgpbd_mg_probe_level0):
    "warm", `reps` launches of each kernel back to back (per-launch time, as the V-cycle runs
    them); "flushed", events around each kernel with a flush_gib GiB buffer rewritten before
    each sweep (larger than the 256 MB last-level cache: the sweep then reads from HBM)."""
    import torch
    from controllable_pbd_3d_amd import _lib
    mg = s._gpu_solver
    n = s.n_cons
    dev = s.device
    g = torch.Generator(device="cpu").manual_seed(1)
    b = torch.randn(n, generator=g).to(dev)
    x = torch.randn(n, generator=g).to(dev)
    y = torch.zeros(n, dtype=torch.float32, device=dev)
    out = {}
    flush = torch.ones(int(flush_gib * (1 << 28)), dtype=torch.float32, device=dev)
    for tag, fb in (("warm", None), ("flushed", flush)):
        ms = np.zeros(3, np.float64)
        mg.ctx("mgpbd_mg_probe_level0", reps, b.data_ptr(), x.data_ptr(), y.data_ptr(),
               None if fb is None else fb.data_ptr(), 0 if fb is None else fb.numel(),
               ms.ctypes.data_as(_lib.F64P))
        out[tag] = {"first_us": ms[0] * 1e3, "second_us": ms[1] * 1e3, "sweep_us": ms[2] * 1e3}
    del flush
    torch.cuda.empty_cache()
    return out


def run_config(mesh, mu_t, lam_t, cfg, name, steps, warmup, profile_steps, model=None, seed=0):
    import torch
    from controllable_pbd_3d_amd.runtime import get_context
    ctx = get_context(0)
    kw = run_kwargs(name, cfg)
    t0 = time.perf_counter()
    s = make_solver(mesh, mu_t, lam_t, cfg, kw,
                    model=model if kw.get("warm_start") == "learned" else None)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"[bench] {name}: setup {setup_s:.1f} s, levels {s._gpu_solver._level_sizes}",
          file=sys.stderr, flush=True)
    load = Loading(cfg, mesh, seed, warmup + steps + profile_steps)
    for k in range(warmup):
        load.step(s)
    torch.cuda.synchronize()
    iters, newton, solves, starts, first_its, predicted = [], [], [], [], [], 0
    t0 = time.perf_counter()
    for k in range(steps):
        load.step(s)
        iters.append(sum(c for _, _, c, _ in s.last_stats))
        newton.append(len(s.last_stats))
        solves.extend(s.last_solve_info)
        starts.append(s.first_solve_info.get("start", 0))
        first_its.append(s.first_solve_info.get("cg_iters", 0))
        predicted += int(getattr(s, "last_predicted", False))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    rel = np.array([np.sqrt(max(d["rr_true"], 0.0) / d["bb"]) if d["bb"] > 0 else 0.0
                    for d in solves])
    conv = np.array([d["reason"] == 1 for d in solves])
    print(f"[bench] {name}: {steps / elapsed:.3f} substeps/s, {np.mean(iters):.1f} PCG it/substep, "
          f"converged {conv.mean() * 100:.0f}% of solves, FP64 rel residual max {rel.max():.2e}",
          file=sys.stderr, flush=True)
    res = {"value": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps, "setup_s": setup_s,
           "n_rows": s.n_cons, "nnz_A0": int(s._nnz),
           "pcg_iters_per_substep": float(np.mean(iters)),
           "newton_iters_per_substep": float(np.mean(newton)),
           "first_solve_pcg_iters": float(np.mean(first_its)),
           "solves": len(solves), "solves_converged": float(conv.mean()) if len(conv) else None,
           "fp64_rel_residual_max": float(rel.max()) if len(rel) else None,
           "fp64_rel_residual_median": float(np.median(rel)) if len(rel) else None}
    res["hbm_footprint_bytes"] = s.device_bytes()
    res["pcg_rescues"], res["pcg_restarts"] = s._gpu_solver.pcg_counters()
    res["pcg_lag_rescues"] = s._gpu_solver.lag_rescues()
    ns, refac = s._gpu_solver.coarse_stats()
    res["coarsest_inverse"] = {"newton_schulz_refresh_steps": ns, "full_inversions": refac}
    if kw.get("warm_start"):
        # start of each substep's first solve: 0 zero, 1 the guess, 2 the previous lambda
        res["first_solve_starts"] = np.bincount(starts, minlength=3).tolist()
    if kw.get("warm_start") == "learned":
        res["network_inferences"] = predicted
    if kw.get("smoother") == "chebyshev":
        res["spectral_estimates"] = [float(v) for v in s._gpu_solver.smoother_estimates()]

    # Kernel durations: HIP events around each launch of the main kernels, on the stream they
    # run on, over `profile_steps` further substeps (the event-bracketed PCG launches eagerly
    # instead of replaying its graph; kernel durations are unchanged, only gaps differ).
    ctx("mgpbd_ctx_profile", 1)
    s._gpu_solver.ctx("mgpbd_ctx_profile", 1)
    for k in range(profile_steps):
        load.step(s)
    cheb = kw.get("smoother") == "chebyshev"
    win0 = any(w["level"] == 0 and w["woff"] is not None for w in s._gpu_solver._win)
    per = 6 if win0 else 8
    fnv = s.n_verts if kw.get("level0") == "factored" else None
    offs = res["slot_offsets"] = fnv is not None and "tbase" in s._gpu_solver._fac
    if offs:    # 0: per-slice row counts; R: every 16-vertex slice padded to R rows
        res["vertex_rows_uniform"] = int(s._gpu_solver._fac["rows_uniform"])
    pk = res["packed_rows"] = fnv is not None and "rows_packed" in s._gpu_solver._fac
    kern = kernel_stats(s._gpu_solver.ctx, s.n_cons, s._nnz, kw["precision"], cheb, per, fnv,
                        offs, pk)
    kern.update({k: v for k, v in kernel_stats(ctx, s.n_cons, s._nnz, kw["precision"], cheb,
                                               per, fnv, offs, pk).items() if k not in kern})
    ctx("mgpbd_ctx_profile", 0)
    s._gpu_solver.ctx("mgpbd_ctx_profile", 0)
    # A rate above the HBM peak means the 256 MB last-level cache (MALL) served part of the
    # kernel's bytes: the smoothing steps of a phase re-read the same level back to back. The
    # line keeps the algorithmic rate and says so; it is not a roofline figure.
    for k, v in kern.items():
        if v.get("GBs", 0.0) > HBM_PEAK_GBS:
            v["above_hbm_peak"] = "part of the bytes served from the 256 MB last-level cache"
            print(f"[bench] {name}: {k} at {v['GBs']:.0f} GB/s of algorithmic bytes, above the "
                  f"{HBM_PEAK_GBS:.0f} GB/s HBM peak (cache-served in part)", file=sys.stderr,
                  flush=True)
    res["kernels"] = kern
    return s, res


def cpu_baseline(mesh, mu_t, lam_t, cfg, omegas, newton_ref, pcg_ref, volume=False):
    """CPU baseline: the oracle's OpenMP C restatement (FP32 arithmetic) of the solver, on the
    GPU host's cores, same mesh, material, hierarchy and PCG tolerance. Converging
    configuration: weighted Jacobi 2 + 2 with fixed per-level weights (from the GPU run's
    spectral estimates), 50 Jacobi sweeps on the coarsest level, relative 1e-6. Timed: one
    full substep on all threads (measured, not scaled); then, on 1 thread, one assembly and
    one PCG iteration, scaled by the all-thread substep's own Newton and PCG counts; and the
    reference algorithm's cost (omega 2/3 on the two-level hierarchy, 50 coarse sweeps), one
    PCG iteration's passes timed on all threads and scaled by the GPU reference run's PCG
    count per substep (it does not converge, so its counts are the caps)."""
    import scipy.sparse as sp
    from oracle import mgpbd_oracle as om
    L, R, _ = om.lib("f32")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L.orc_set_threads(threads)
    args = solver_args(mesh, mu_t, lam_t, volume=volume)
    okeys = ("v_p", "v_p_ref", "v_invm", "t_i", "t_m", "gravity", "dt", "maxiter", "atol",
             "rtol", "maxiter_cg", "tol_cg", "relative_tol", "n_smooth", "omega_jacobi",
             "use_line_search", "mu_per_tet", "hold_pinned", "volume_lambda")
    o = om.OracleMGPBD(**{k: v for k, v in args.items() if k in okeys}, precision="f32",
                       omega_levels=omegas)
    from controllable_pbd_3d_amd import tet_aggregation
    aggs = tet_aggregation.build_extended_hierarchy(mesh.tets, mesh.n_tets, max_coarse=64,
                                                    max_levels=cfg["max_levels"] - 1)
    Ps = [sp.csr_matrix((np.ones(len(a)), (np.arange(len(a)), a)), shape=(len(a), n))
          for a, n in aggs]
    t0 = time.perf_counter()
    o.init(Ps=Ps, structural=True)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.step(0)
    sub_s = time.perf_counter() - t0
    hist = o.history[-1]
    newton = len(hist)
    pcg = sum(c for _, _, c, _ in hist)

    def assembly():
        t0 = time.perf_counter()
        o._compute_C_and_gradC()
        o._fill_A(o.A_data)
        o.mg.update_fine_matrix_data(o.A_data)
        o.mg.update_all_coarse_matrices()
        return time.perf_counter() - t0

    def pcg_iteration():
        o.mg.maxiter_cg = 1
        o.dlam[:] = 0
        t0 = time.perf_counter()
        o.mg.solve(o.b, o.dlam)
        return time.perf_counter() - t0

    t_asm_n = assembly()
    L.orc_set_threads(1)
    t_asm1, t_it1 = assembly(), pcg_iteration()
    one_thread = 1.0 / (newton * t_asm1 + pcg * t_it1)
    # the reference algorithm's PCG iteration on all threads: level 0 SpMV + 2 + 2 sweeps +
    # residual (6 level-0 passes) and 50 Jacobi sweeps of its level 1 (this hierarchy's level
    # 1: the reference's tet aggregation), 100 level-1 passes
    L.orc_set_threads(threads)
    lv0, lv1 = o.mg.levels[0], o.mg.levels[1]
    v0 = np.zeros(lv0["n"], np.float32)
    v1 = np.zeros(lv1["n"], np.float32)
    t0 = time.perf_counter()
    for _ in range(6):
        L.orc_spmv(lv0["n"], om.ptr(lv0["ip"], om.I32P), om.ptr(lv0["ix"], om.I32P),
                   om.ptr(lv0["a"], R), om.ptr(lv0["b"], R), om.ptr(v0, R))
    for _ in range(100):
        L.orc_spmv(lv1["n"], om.ptr(lv1["ip"], om.I32P), om.ptr(lv1["ix"], om.I32P),
                   om.ptr(lv1["a"], R), om.ptr(lv1["b"], R), om.ptr(v1, R))
    t_ref_it = time.perf_counter() - t0
    cpu_model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": 1.0 / sub_s, "unit": "substeps/s", "cores": threads, "kind": "port",
            "sample": (f"one full substep of the 1st frame on the same mesh ({mesh.n_tets} tets"
                       f"{', volume constraint' if volume else ''}), FP32 OpenMP C oracle, "
                       f"{threads} threads: {sub_s:.1f} s, {newton} Newton / {pcg} PCG "
                       f"iterations to relative 1e-6 (weighted Jacobi 2+2 with per-level "
                       f"weights {np.round(omegas, 3).tolist()}, same 5 levels, 50 coarsest "
                       f"sweeps); setup {setup_s:.1f} s excluded"),
            "one_thread_substeps_per_s": one_thread,
            "one_thread_sample": (f"1 thread: assembly {t_asm1:.2f} s, PCG iteration "
                                  f"{t_it1:.2f} s, scaled by the substep's {newton} Newton and "
                                  f"{pcg} PCG iterations"),
            "reference_algorithm_substeps_per_s": 1.0 / (newton_ref * t_asm_n +
                                                         pcg_ref * t_ref_it),
            "reference_algorithm_sample": (
                f"{threads} threads: assembly {t_asm_n:.2f} s, one PCG iteration of the "
                f"reference algorithm (6 level-0 + 100 level-1 passes) {t_ref_it:.2f} s, scaled "
                f"by the GPU reference run's {newton_ref:.1f} Newton and {pcg_ref:.0f} PCG "
                f"iterations per substep (capped at 100 per solve, unconverged)"),
            "nproc": os.cpu_count(), "cpu_model": cpu_model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS),
                    help="c1 (default) 1M-tet block, E 1e4-1e6; c2 the same block with E "
                         "1e4-1e8; c3 196,608-tet block with drag loading; c4 ~6M-tet "
                         "ellipsoid with a stiff inclusion and drag loading; c0 3,072 tets")
    ap.add_argument("--seed", type=int, default=0,
                    help="mesh jitter / material / drag seed of the measured workload")
    ap.add_argument("--runs", default="all",
                    help="comma list of " + ",".join(RUNS) + " (default: all)")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--probe-reps", type=int, default=60)
    ap.add_argument("--train-seeds", default="1,2")
    ap.add_argument("--train-substeps", type=int, default=8)
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--python-loop", action="store_true",
                    help="run the Newton loop in Python instead of mgpbd_substep")
    args = ap.parse_args()
    global NATIVE_SUBSTEP
    NATIVE_SUBSTEP = not args.python_loop
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    cfg = CONFIGS[args.config]
    names = DEFAULT_RUNS if args.runs == "all" else args.runs.split(",")
    for n in names:
        if n not in RUNS:
            raise SystemExit(f"unknown run {n}")
    print(f"[bench] building {cfg['label']}", file=sys.stderr, flush=True)
    mesh, mu_t, lam_t = build_workload(cfg, args.seed)
    torch.cuda.set_device(0)
    model, learn_info = None, None
    if "fp16_learned" in names:
        model, learn_info = train_learned(cfg, [int(x) for x in args.train_seeds.split(",")],
                                          args.train_substeps, args.train_epochs)
    head = "fp16_learned" if "fp16_learned" in names else names[-1]
    runs, pos, probe, omegas, keep = {}, {}, None, None, None
    for name in names:
        s, res = run_config(mesh, mu_t, lam_t, cfg, name, args.steps, args.warmup,
                            args.profile_steps, model, args.seed)
        runs[name] = res
        pos[name] = s.v_p.detach().cpu().numpy().astype(np.float64)
        if name == head:
            probe = probe_level0(s, args.probe_reps)
            res["level0_probe"] = probe
        if name == "fp32":
            # per-level Jacobi weights for the CPU baseline's converging configuration
            omegas = [float((4.0 / 3.0) / (1.2 * max(l, 1.0))) for l in res["spectral_estimates"]]
        if keep is None:
            keep = (s._gpu_solver._level_sizes, int(s._nnz))
        del s
        torch.cuda.empty_cache()
    for group in (("fp32", "fp32_reference", "fp16_assembled", "fp16", "bf16", "fp16_previous",
                   "fp16_learned"), ("fp32_volume", "fp16_volume")):
        # every run of a group takes the same substeps under the same loading: positions after
        # the last one against the group's FP32 run (FP64-verified solves), in edge lengths
        ref = group[0]
        if ref in pos:
            for name in group:
                if name in pos:
                    d = np.abs(pos[name] - pos[ref]).max()
                    runs[name][f"max_position_diff_vs_{ref}_over_h"] = float(d / mesh.h)
    best = runs[head]
    prec = RUNS[head]["precision"]
    factored = RUNS[head].get("level0") == "factored"
    out = {
        "metric": "MGPBD substeps/s at 1M tets to 1e-6 relative residual (FP64-verified): "
                  "mixed-precision V-cycle + learned warm start vs plain FP32 GPU MGPCG"
                  if args.config == "c1" else f"MGPBD substeps/s on {cfg['label']} to 1e-6 "
                  "relative residual (FP64-verified)",
        "value": best["value"], "unit": "substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else (
            f"{prec} V-cycle (f32 accumulation), f64 PCG products and dots (factored level 0)"
            if factored else f"f32 outer PCG (f64 dots), {prec} V-cycle"),
        "data": "synthetic",
        "config": {"workload": cfg["label"], "seed": args.seed, "drag": bool(cfg.get("drag")),
                   "n_tets": mesh.n_tets, "n_verts": mesh.n_verts,
                   "nnz_A0": keep[1], "levels": keep[0],
                   "pcg_rel_tol": 1e-6, "pcg_maxiter": cfg.get("maxiter_cg", 100),
                   "E_range": cfg["E"], "nu": 0.45, "reported": head,
                   "newton_loop": "mgpbd_substep (C++)" if NATIVE_SUBSTEP else "Python",
                   "runs": runs, "learned_training": learn_info,
                   "speedup_vs_fp32": (best["value"] / runs["fp32"]["value"]
                                       if "fp32" in runs else None),
                   "speedup_fp16_assembled_vs_fp32": (
                       runs["fp16_assembled"]["value"] / runs["fp32"]["value"]
                       if "fp16_assembled" in runs and "fp32" in runs else None),
                   "speedup_fp16_volume_vs_fp32_volume": (
                       runs["fp16_volume"]["value"] / runs["fp32_volume"]["value"]
                       if "fp16_volume" in runs and "fp32_volume" in runs else None)},
    }
    if probe is not None and factored:
        # dominant kernel of the reported run: the factored level 0's vertex pass
        # (k_fac_gather), `reps` launches back to back between two HIP events
        n_rows = best["n_rows"]
        by = factored_bytes(n_rows, mesh.n_verts, "gather", offsets=best["slot_offsets"])
        us = probe["warm"]["first_us"]
        gbs = by / (us * 1e-6) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": f"k_fac_gather ({head})", "achieved": gbs,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                           "algorithmic_bytes": by, "avg_us": us,
                           "traffic": measured_traffic(cfg["label"], "k_fac_gather"),
                           "apply_pass": {"kernel": "k_fac_apply<1>",
                                          "avg_us": probe["warm"]["second_us"],
                                          "GBs": factored_bytes(n_rows, mesh.n_verts, "apply",
                                                                packed=best["packed_rows"]) /
                                          (probe["warm"]["second_us"] * 1e-6) / 1e9},
                           "cache_flushed": probe["flushed"]}
    if args.cpu_baseline:
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        ref = runs.get("fp32_reference", best)
        if omegas is None:
            omegas = [0.2, 0.5]
        out["cpu_baseline"] = cpu_baseline(mesh, mu_t, lam_t, cfg, omegas,
                                           ref["newton_iters_per_substep"],
                                           ref["pcg_iters_per_substep"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
