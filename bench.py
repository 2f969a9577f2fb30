"""Benchmark: MGPBD substeps/s on one MI355X.

Default workload (the 1M-tet configuration): jittered block of 56^3 Kuhn cubes (185,193
vertices, 1,053,696 tets), h = 0.02 m, per-tet E log-uniform in [1e4, 1e6] Pa (seed 0),
nu = 0.45, rho = 1000 kg/m^3, gravity -9.8 m/s^2, top face held, dt = 1/300 s (60 fps x 5
substeps), Newton defaults of the reference (maxiter 20, atol 1e-4, rtol 1e-2, line search),
PCG to relative residual 1e-6 (max 100 iterations), 2+2 weighted-Jacobi sweeps (omega 2/3),
tet-aggregation hierarchy extended to 5 levels. A step is one substep (predict, Newton/MGPCG
solve, velocity update) from the state the previous step left; inputs are resident in HBM.

Prints one JSON line (see README for the fields). `--gpus N` with N > 1 exits: this build
measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (cubes, h, E range, levels)
    "c1": dict(n=56, h=0.02, E=(1e4, 1e6), max_levels=5, label="box56_1053696tets"),
    "c3": dict(n=32, h=0.03, E=(1e4, 1e6), max_levels=4, label="box32_196608tets"),
    "c0": dict(n=8, h=0.05, E=None, max_levels=3, label="box8_3072tets"),
}


def build_workload(cfg, seed=0):
    from controllable_pbd_3d_amd.mesh import box_mesh
    mesh = box_mesh(cfg["n"], cfg["h"], jitter=0.05, seed=seed, rho=1000.0,
                    E_range=cfg["E"] if cfg["E"] else (1e5, 1e5))
    mu_t = (mesh.E / (2.0 * (1.0 + 0.45))).astype(np.float32)
    return mesh, mu_t


def solver_args(mesh, mu_t):
    return dict(v_p=mesh.verts.copy(), v_p_ref=mesh.verts.copy(), v_invm=mesh.invm.copy(),
                t_i=mesh.tets.copy(), t_m=mesh.tet_mass.copy(), gravity=(0.0, -9.8, 0.0),
                dt=1.0 / 60.0 / 5.0, maxiter=20, atol=1e-4, rtol=1e-2, maxiter_cg=100,
                n_smooth=2, omega_jacobi=2.0 / 3.0, use_line_search=True, mu_per_tet=mu_t)


def kernel_bytes(n_rows, nnz, kind):
    """Algorithmic HBM bytes per launch (FP32 values, int32 columns; each array once)."""
    if kind == "spmv":      # values + columns + row pointers + x + y
        return 8 * nnz + 4 * (n_rows + 1) + 4 * n_rows + 4 * n_rows
    if kind == "jacobi":    # + b, D^-1, x (own) -> x_new
        return 8 * nnz + 4 * (n_rows + 1) + 16 * n_rows
    if kind == "resid":     # values + columns + ptr + x + b -> r
        return 8 * nnz + 4 * (n_rows + 1) + 12 * n_rows
    raise ValueError(kind)


def cpu_baseline(mesh, mu_t, budget_s=30.0):
    """The oracle's CPU restatement of the reference algorithm (two-level tet hierarchy,
    absolute PCG tolerance 1e-5, 50 coarse sweeps, line search; FP32, one thread) on the same
    mesh and material, timed per substep after its setup."""
    from oracle import mgpbd_oracle as om
    args = solver_args(mesh, mu_t)
    o = om.OracleMGPBD(**args, precision="f32", hold_pinned=True)
    t0 = time.perf_counter()
    o.init()
    setup_s = time.perf_counter() - t0
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        o.step(0)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 5:
            break
    per = float(np.median(times))
    return {"value": 1.0 / per, "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": f"{len(times)} substep(s) of the same mesh from rest, reference algorithm "
                      f"(2-level tet hierarchy, abs tol 1e-5), FP32 C oracle, 1 thread; "
                      f"setup {setup_s:.1f} s excluded; host cores available {os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    from controllable_pbd_3d_amd.runtime import get_context
    from controllable_pbd_3d_amd import _lib

    cfg = CONFIGS[args.config]
    print(f"[bench] building {cfg['label']}", file=sys.stderr, flush=True)
    mesh, mu_t = build_workload(cfg)
    torch.cuda.set_device(0)
    ctx = get_context(0)
    t0 = time.perf_counter()
    s = MGPBDSolverGPU(**solver_args(mesh, mu_t), hold_pinned=True, relative_tol=True,
                       tol_cg=1e-6, hierarchy="extended", galerkin="structural",
                       max_coarse=64, max_levels=cfg["max_levels"], verbose=False)
    s.init()
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    levels = s._gpu_solver._level_sizes
    print(f"[bench] setup {setup_s:.1f} s, levels {levels}, nnz {s._nnz}", file=sys.stderr,
          flush=True)
    for k in range(args.warmup):
        s.step(0)
    torch.cuda.synchronize()
    ctx("mgpbd_ctx_profile", 1)
    iters = []
    newton = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step(0)
        iters.append(sum(c for _, _, c, _ in s.last_stats))
        newton.append(len(s.last_stats))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = np.zeros(8, np.float64)
    cnt = np.zeros(8, np.int64)
    ctx("mgpbd_ctx_profile_read", ms.ctypes.data_as(_lib.F64P),
        cnt.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_int64)))
    ctx("mgpbd_ctx_profile", 0)
    n, nnz = s.n_tets, s._nnz
    kern = {}
    for kind, slot in (("spmv", 0), ("jacobi", 1), ("resid", 2)):
        if cnt[slot] > 0:
            avg = ms[slot] / cnt[slot] * 1e-3
            by = kernel_bytes(n, nnz, kind)
            kern[kind] = {"avg_us": avg * 1e6, "launches": int(cnt[slot]),
                          "total_ms": float(ms[slot]), "GBs": by / avg / 1e9, "bytes": by}
    dom = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    value = args.steps / elapsed
    out = {
        "metric": "MGPBD substeps/s at 1M tets to 1e-6 residual (plain FP32 HIP MGPCG)",
        "value": value, "unit": "substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["label"], "n_tets": n, "n_verts": s.n_verts, "nnz_A0": nnz,
                   "levels": levels, "precision": "fp32", "pcg_rel_tol": 1e-6,
                   "E_range": cfg["E"], "setup_s": setup_s,
                   "pcg_iters_per_substep": float(np.mean(iters)),
                   "newton_iters_per_substep": float(np.mean(newton)),
                   "kernels": kern},
    }
    if dom:
        k = kern[dom]
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": k["GBs"],
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": k["GBs"] / HBM_PEAK_GBS, "traffic": None}
    if args.cpu_baseline:
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(mesh, mu_t, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
