"""Experiment drivers on the GPU solver: resolution scaling, stiffness sensitivity, ablation.

They play the role of the reference's experiments/exp_resolution.py (1.2),
experiments/exp_stiffness.py (1.3) and experiments/exp_ablation.py (3.1): every run records
its frames with ``SolverBenchmark`` (one JSON document per run, the reference's format) and the
driver writes the list of run summaries its plot helper takes
(experiments/plot_utils.py: ``plot_resolution_scaling`` reads solver_name, n_tets,
avg_frame_time and fps; ``plot_stiffness_scaling`` solver_name, mu, avg_iterations and
avg_frame_time; ``plot_timing_bars`` solver_name and avg_frame_time).

Meshes: the synthetic jittered Kuhn blocks (top face held) at several resolutions, and the
reference's Medit meshes from a directory given by ``--mesh-dir`` (simple_sphere, spot,
cube_tet; vertex 0 held). Solvers compared on each mesh (all ``MGPBDSolverGPU``):

* ``MGPBD-GPU reference``: the reference's algorithm (two tet-aggregation levels, weighted
  Jacobi 2 + 2 with omega 2/3, 50 coarsest sweeps, absolute PCG tolerance 1e-5);
* ``MGPBD-GPU extended``: deeper hierarchy, Chebyshev smoothing on spectral estimates, dense
  coarsest solve, FP64-verified relative 1e-6 PCG;
* ``MGPBD-GPU extended FP16``: the same with the FP16 V-cycle on the factored level 0.
The ablation adds the pyamg-style standard aggregation and the adaptive zone.

  python -m controllable_pbd_3d_amd.experiments resolution --out out/experiments/resolution
  python -m controllable_pbd_3d_amd.experiments stiffness --out out/experiments/stiffness
  python -m controllable_pbd_3d_amd.experiments ablation --out out/experiments/ablation
"""

from __future__ import annotations

import argparse
import json
import os
import sys

from .benchmark import SolverBenchmark
from .mesh import box_mesh, load_medit_mesh

FPS, SUBSTEPS = 60, 5
REFERENCE = dict(mu=1e6, maxiter=20, atol=1e-4, rtol=1e-2, maxiter_cg=100, tol_cg=1e-5,
                 n_smooth=2, omega_jacobi=2.0 / 3.0, damp=0.993)
EXTENDED = dict(hierarchy="extended", galerkin="structural", max_coarse=64,
                coarse_solver="cholesky", relative_tol=True, tol_cg=1e-6, smoother="chebyshev",
                n_smooth=5, smooth_schedule=[5, 1], cheb_ratio=30.0, smoother_safety=1.2,
                verify_residual=1, fp64_iterate=True, hold_pinned=True)
VARIANTS = {
    "MGPBD-GPU reference": {},
    "MGPBD-GPU extended": dict(EXTENDED),
    "MGPBD-GPU extended FP16": dict(EXTENDED, precision="fp16", level0="factored"),
}


def block(n: int, h: float = 0.05, seed: int = 0):
    m = box_mesh(n, h, jitter=0.05, seed=seed)
    m.name = f"box{n}"
    return m


def medit(mesh_dir: str, name: str):
    path = os.path.join(mesh_dir, name + ".mesh")
    if not os.path.exists(path):
        path = os.path.join(mesh_dir, name, name + ".mesh")
    if not os.path.exists(path):
        return None
    m = load_medit_mesh(path, fixed_ids=[0])
    m.name = name
    return m


def run(mesh, solver_name: str, kw: dict, n_frames: int, gravity=(0.0, -9.8, 0.0),
        adaptive_radius=None, bench_dir=None, extra=None) -> dict:
    """Runs n_frames frames of SUBSTEPS substeps on ``mesh``; returns the benchmark summary
    with n_tets and fps (and ``extra``), saving the frame record when bench_dir is given."""
    from .mgpbd_gpu import MGPBDSolverGPU
    params = dict(REFERENCE)
    params.update(kw)
    bench = SolverBenchmark(solver_name, mesh.name, dict(extra or {}, n_tets=mesh.n_tets))
    s = MGPBDSolverGPU(v_p=mesh.verts.copy(), v_p_ref=mesh.verts.copy(),
                       v_invm=mesh.invm.copy(), t_i=mesh.tets.copy(), t_m=mesh.tet_mass.copy(),
                       gravity=gravity, dt=1.0 / FPS / SUBSTEPS, use_line_search=True,
                       benchmark=bench, verbose=False, **params)
    s.init()
    stepper = s
    if adaptive_radius is not None:
        from .adaptive import AdaptiveZone, MGPBDAdaptiveSolver
        zone = AdaptiveZone(s.n_tets, s.t_i, s.v_p)
        stepper = MGPBDAdaptiveSolver(s, zone)
    centre = mesh.verts.mean(axis=0)
    for frame in range(n_frames):
        for _ in range(SUBSTEPS):
            if adaptive_radius is not None:
                stepper.step(frame, tool_pos=centre, tool_radius=adaptive_radius)
            else:
                stepper.step(frame)
    out = bench.summary()
    out["n_tets"] = mesh.n_tets
    out["fps"] = 1.0 / out["avg_frame_time"] if out["avg_frame_time"] > 0 else 0.0
    out.update(extra or {})
    if bench_dir:
        os.makedirs(bench_dir, exist_ok=True)
        label = f"{solver_name}_{mesh.name}" + "".join(f"_{k}{v:g}" for k, v in (extra or {}).items())
        tag = "".join(c if c.isalnum() else "_" for c in label)
        bench.save(os.path.join(bench_dir, tag + ".json"))
    return out


def resolution(args) -> list:
    meshes = [block(n) for n in args.sizes]
    for name in ("simple_sphere", "spot", "cube_tet"):
        m = medit(args.mesh_dir, name) if args.mesh_dir else None
        if m is not None:
            meshes.append(m)
    results = []
    for m in meshes:
        for name, kw in VARIANTS.items():
            r = run(m, name, kw, args.frames, bench_dir=os.path.join(args.out, "runs"))
            print(f"[resolution] {name} {m.name} ({m.n_tets} tets): "
                  f"{r['avg_frame_time'] * 1e3:.2f} ms/frame", flush=True)
            results.append(r)
    return results


def stiffness(args) -> list:
    m = block(args.sizes[-1])
    results = []
    for mu in args.mus:
        for name, kw in VARIANTS.items():
            r = run(m, name, dict(kw, mu=mu), args.frames,
                    bench_dir=os.path.join(args.out, "runs"), extra={"mu": mu})
            print(f"[stiffness] {name} mu={mu:.0e}: {r['avg_iterations']:.1f} Newton "
                  f"iterations/frame, {r['avg_frame_time'] * 1e3:.2f} ms/frame", flush=True)
            results.append(r)
    return results


def ablation(args) -> list:
    m = block(args.sizes[-1])
    steps = [("(a) reference algorithm", {}, None),
             ("(b) +pyamg-style aggregation", dict(aggregation="pyamg"), None),
             ("(c) +extended hierarchy, Chebyshev, dense coarsest", dict(EXTENDED), None),
             ("(d) +FP16 factored level 0", VARIANTS["MGPBD-GPU extended FP16"], None),
             ("(e) +adaptive zone", VARIANTS["MGPBD-GPU extended FP16"], 0.3 * m.h * args.sizes[-1])]
    results = []
    for name, kw, radius in steps:
        r = run(m, name, kw, args.frames, adaptive_radius=radius,
                bench_dir=os.path.join(args.out, "runs"))
        results.append(r)
    base = results[0]["avg_frame_time"]
    for r in results:
        r["speedup"] = base / r["avg_frame_time"] if r["avg_frame_time"] > 0 else 0.0
        print(f"[ablation] {r['solver_name']}: {r['avg_frame_time'] * 1e3:.2f} ms/frame "
              f"({r['speedup']:.2f}x)", flush=True)
    return results


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("experiment", choices=("resolution", "stiffness", "ablation"))
    ap.add_argument("--out", default="out/experiments")
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--sizes", type=int, nargs="+", default=[4, 8, 16])
    ap.add_argument("--mus", type=float, nargs="+", default=[1e3, 1e4, 1e5, 1e6, 1e7])
    ap.add_argument("--mesh-dir", default=None,
                    help="directory holding the reference's Medit meshes (optional)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    results = {"resolution": resolution, "stiffness": stiffness, "ablation": ablation}[
        args.experiment](args)
    path = os.path.join(args.out, f"{args.experiment}.json")
    with open(path, "w") as fh:
        json.dump(results, fh, indent=2)
    print(json.dumps({"experiment": args.experiment, "runs": len(results), "out": path}))
    return results


if __name__ == "__main__":
    sys.exit(0 if main() is not None else 1)
