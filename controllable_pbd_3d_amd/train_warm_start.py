"""Train and evaluate the learned warm start on the spring-drag workload.

Workload: a jittered block of n^3 Kuhn cubes (6 tets each), per-tet E log-uniform in
[1e4, 1e6] Pa (nu = 0.45, rho = 1000), top face held; every frame a seeded random free vertex
is pulled by a spring (k = 1e3 N/m) toward a target uniform in a 0.5 m cube around its rest
position; 5 substeps per frame at dt = 1/300 s; PCG to a relative residual of 1e-6, every solve
FP64-verified; Chebyshev smoother on the spectral estimates, FP16 V-cycle.

Training pairs come from the solver itself: the first Newton iteration of every substep on the
training seeds (the solver runs from the previous lambda meanwhile). The network is trained on
the GPU with the A-norm loss (x0 - x*)^T A (x0 - x*) / x*^T A x* of the start x0 against the
converged dlambda x* (the quantity that bounds CG's iterations). Evaluation on a held-out seed
solves every substep's first system twice, from the previous lambda and from the learned start
(the device keeps whichever of prediction / previous lambda has the smaller residual), and
also runs the whole sequence once per start to time it, under the strong preconditioner the
solver uses and under the reference's weak V-cycle (PRECONDITIONERS).

  python -m controllable_pbd_3d_amd.train_warm_start --n 32 --seeds 0-3 --frames 25 \\
      --eval-seed 100 --eval-frames 10 --out warm_start_box32.npz
prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from .mesh import box_mesh, drag_schedule
from .mgpbd_gpu import MGPBDSolverGPU
from .warm_start import LearnedWarmStart


def workload(n: int, h: float, seed: int):
    mesh = box_mesh(n, h, jitter=0.05, seed=seed, rho=1000.0, E_range=(1e4, 1e6))
    mu_t = (mesh.E / (2.0 * 1.45)).astype(np.float32)
    return mesh, mu_t


# preconditioners the start is evaluated under (the solution, and so the training target, is
# the same for both): "chebyshev", the extended hierarchy with Chebyshev smoothing and the
# dense coarsest solve (each PCG iteration contracts the error ~20x or more); "two_level", the
# reference's hierarchy (tet aggregation, two levels, 50 coarsest sweeps, FP32) with 2 + 2
# weighted-Jacobi sweeps whose weights come from the spectral estimates (the reference's fixed
# 2/3 does not contract on every level of these systems), a weak preconditioner that needs
# tens of iterations per solve
PRECONDITIONERS = {
    "chebyshev": dict(n_smooth=5, hierarchy="extended", galerkin="structural", max_coarse=64,
                      coarse_solver="cholesky", smoother="chebyshev", cheb_ratio=30.0,
                      smoother_safety=1.2, smooth_schedule=[5, 1], maxiter_cg=100),
    "two_level": dict(n_smooth=2, hierarchy="reference", galerkin="values",
                      coarse_solver="jacobi", smoother="spectral", precision="fp32",
                      maxiter_cg=1000),
}


def make_solver(mesh, mu_t, max_levels, warm_start="none", model=None, record=None,
                solver_cls=MGPBDSolverGPU, precision="fp16", precond="chebyshev"):
    kw = dict(max_levels=max_levels, precision=precision)
    kw.update(PRECONDITIONERS[precond])
    if precond == "two_level":
        kw.pop("max_levels")
    s = solver_cls(v_p=mesh.verts.copy(), v_p_ref=mesh.verts.copy(), v_invm=mesh.invm.copy(),
                   t_i=mesh.tets.copy(), t_m=mesh.tet_mass.copy(), gravity=(0.0, -9.8, 0.0),
                   dt=1.0 / 300.0, maxiter=20, atol=1e-4, rtol=1e-2,
                   omega_jacobi=2.0 / 3.0, mu_per_tet=mu_t, hold_pinned=True,
                   relative_tol=True, tol_cg=1e-6, verify_residual=1, fp64_iterate=True,
                   warm_start=warm_start, warm_start_model=model, record_samples=record,
                   verbose=False, **kw)
    s.init()
    return s


def run_sequence(s, mesh, seed, frames, k_spring=1e3, box=0.5, per_substep=None):
    for f, (v, target) in enumerate(drag_schedule(mesh, seed, frames, box)):
        s.set_drag(v, target, k_spring)
        for _ in range(5):
            s.step(f)
            if per_substep is not None:
                per_substep(s)


class _Shadow(MGPBDSolverGPU):
    """Solves every substep's first system from the previous lambda as well (not timed)."""

    def _solve_warm_started(self):
        b = self._b_gpu
        x = self._lam_prev.clone()
        it_prev = self._gpu_solver.solve_guess(b, x)
        info_prev = dict(self._gpu_solver.last_info)
        it = super()._solve_warm_started()
        info = dict(self._gpu_solver.last_info)
        self.shadow.append({"prev_iters": it_prev, "learned_iters": it,
                            "prev_rr0": info_prev["rr0"] / info_prev["bb"],
                            "learned_rr_guess": info["rr_guess"] / info["bb"],
                            "start": info["start"]})
        return it


def evaluate(model, n, h, max_levels, seed, frames, warmup_frames=1, box=0.5,
             precond="chebyshev"):
    mesh, mu_t = workload(n, h, seed)
    out = {"preconditioner": precond}
    # 1. the same systems from both starts
    s = make_solver(mesh, mu_t, max_levels, "learned", model, solver_cls=_Shadow,
                    precond=precond)
    s.shadow = []
    run_sequence(s, mesh, seed, frames, box=box)
    sh = s.shadow[5 * warmup_frames:]
    out["paired"] = {
        "substeps": len(sh),
        "first_solve_iters_previous": float(np.mean([d["prev_iters"] for d in sh])),
        "first_solve_iters_learned": float(np.mean([d["learned_iters"] for d in sh])),
        "initial_rel_residual_previous_median": float(np.median([np.sqrt(d["prev_rr0"]) for d in sh])),
        "initial_rel_residual_learned_median": float(np.median([np.sqrt(d["learned_rr_guess"]) for d in sh])),
        "starts_taken": np.bincount([d["start"] for d in sh], minlength=3).tolist(),
    }
    del s
    # 2. whole sequences, timed
    for mode in ("previous", "learned"):
        s = make_solver(mesh, mu_t, max_levels, mode, model if mode == "learned" else None,
                        precond=precond)
        run_sequence(s, mesh, seed, warmup_frames, box=box)
        stats = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_sequence(s, mesh, seed, frames - warmup_frames,
                     box=box, per_substep=lambda z: stats.append(
                         (sum(c for _, _, c, _ in z.last_stats), z.first_solve_info.get("cg_iters", 0),
                          max([np.sqrt(max(d["rr_true"], 0.0) / d["bb"]) for d in z.last_solve_info
                               if d["bb"] > 0], default=0.0))))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k = len(stats)
        out[mode] = {"substeps": k, "ms_per_substep": 1e3 * dt / max(k, 1),
                     "pcg_iters_per_substep": float(np.mean([a for a, _, _ in stats])),
                     "first_solve_iters": float(np.mean([b for _, b, _ in stats])),
                     "fp64_rel_residual_max": float(max(c for _, _, c in stats))}
        del s
    return out


def parse_seeds(text: str):
    if "-" in text:
        a, b = text.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(x) for x in text.split(",")]


def generate_pairs(n, h, max_levels, seeds, frames, model, box=0.5):
    """Training pairs of the first Newton iteration of every substep on the given seeds."""
    samples = []
    for seed in seeds:
        mesh, mu_t = workload(n, h, seed)
        s = make_solver(mesh, mu_t, max_levels, "previous", model, record=samples)
        run_sequence(s, mesh, seed, frames, box=box)
        del s
        print(f"[train] seed {seed}: {len(samples)} pairs", file=sys.stderr, flush=True)
    return samples


def train_model(model, samples, epochs, n_heldout=25):
    """Adam on the A-norm loss; returns (train-loss history, held-out loss before, after).
    The held-out pairs are the last ``n_heldout`` (the last frames of the last seed)."""
    held = samples[-n_heldout:]
    train = samples[:-n_heldout] if len(samples) > 2 * n_heldout else samples
    untrained = model.evaluate(held)
    hist = []
    for e in range(epochs):
        hist += model.train(train, 1, "energy", seed=e)
        print(f"[train] epoch {e}: train loss {hist[-1]:.4e}", file=sys.stderr, flush=True)
    return hist, untrained, model.evaluate(held)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--h", type=float, default=0.03)
    ap.add_argument("--max-levels", type=int, default=4)
    ap.add_argument("--seeds", default="0-3")
    ap.add_argument("--frames", type=int, default=25)
    ap.add_argument("--box", type=float, default=0.5, help="drag target cube side (m)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-seed", type=int, default=100)
    ap.add_argument("--eval-frames", type=int, default=10)
    ap.add_argument("--out", default=None, help="checkpoint (.npz) to write")
    ap.add_argument("--precond", default="chebyshev,two_level",
                    help="comma list of preconditioners to evaluate under: " +
                         ",".join(PRECONDITIONERS))
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    seeds = parse_seeds(args.seeds)
    mesh0, _ = workload(args.n, args.h, seeds[0])
    model = LearnedWarmStart.for_mesh(mesh0.tets, mesh0.n_tets, seed=0, lr=args.lr)
    t0 = time.perf_counter()
    samples = generate_pairs(args.n, args.h, args.max_levels, seeds, args.frames, model, args.box)
    gen_s = time.perf_counter() - t0
    n_pairs = len(samples)
    t1 = time.perf_counter()
    hist, untrained, trained = train_model(model, samples, args.epochs)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    del samples
    torch.cuda.empty_cache()
    if args.out:
        model.save(args.out)
    ev = {p: evaluate(model, args.n, args.h, args.max_levels, args.eval_seed, args.eval_frames,
                      box=args.box, precond=p) for p in args.precond.split(",")}
    out = {"workload": f"box{args.n}_{mesh0.n_tets}tets_drag", "train_seeds": seeds,
           "frames_per_seed": args.frames, "pairs": n_pairs, "drag_box_m": args.box,
           "generate_s": gen_s, "train_s": train_s, "epochs": args.epochs,
           "train_loss": hist, "heldout_energy_loss_untrained": untrained,
           "heldout_energy_loss_trained": trained, "eval_seed": args.eval_seed,
           "evaluation": ev}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
