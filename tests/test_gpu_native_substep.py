"""The library's substep (mgpbd_substep: prediction, Newton loop, line search and velocity update
in C++) against the Python Newton loop of the same solver class (cons/mgpbd_gpu.py:200-219 and
397-446): two solvers built alike, one with native_substep=True, stepped side by side. Positions,
velocities and lambda must agree bit for bit after every substep, and so must the per-iteration
records (r, PCG iterations, line-search step) and the PCG's solve records."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import deformed, physical_block, solver_kwargs   # noqa: E402

EXT = dict(hierarchy="extended", galerkin="structural", max_coarse=64, relative_tol=True,
           tol_cg=1e-6, hold_pinned=True)
MIXED = dict(EXT, coarse_solver="cholesky", smoother="chebyshev", verify_residual=1,
             fp64_iterate=True, rtol=1e-6, atol=1e-9, maxiter=30, spectral_refresh="substep")

CASES = {
    # the reference's solver semantics from a deformed start
    "reference": (dict(), True, False),
    "reference_no_line_search": (dict(use_line_search=False), True, False),
    "extended_fp32_drag": (dict(EXT, smoother="spectral"), False, True),
    "volume_fp32": (dict(EXT, volume=True, omega_levels=[0.25, 2.0 / 3.0]), True, False),
    "fp16_factored_previous_drag": (dict(MIXED, precision="fp16", level0="factored",
                                         warm_start="previous"), False, True),
    "bf16_factored_volume": (dict(MIXED, precision="bf16", level0="factored", volume=True),
                             False, False),
    "fp16_learned_overlap_drag": (dict(MIXED, precision="fp16", level0="factored",
                                       warm_start="learned"), False, True),
    "fp16_learned_serial": (dict(MIXED, precision="fp16", level0="factored",
                                 warm_start="learned", overlap_inference=False), True, False),
    # levels 1.. rebuilt at each substep's first Newton iteration only
    "fp16_factored_coarse_lag_drag": (dict(MIXED, precision="fp16", level0="factored",
                                           coarse_refresh="substep"), True, True),
    "fp32_coarse_lag": (dict(EXT, smoother="spectral", coarse_refresh="substep"), True, False),
    # levels 1.. rebuilt at the first Newton iteration and after solves of >= 3 PCG iterations
    "fp16_factored_coarse_adaptive_drag": (dict(MIXED, precision="fp16", level0="factored",
                                                coarse_refresh="adaptive",
                                                coarse_refresh_iters=3), True, True),
}


def _pair(kw, start_deformed, seed):
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    from controllable_pbd_3d_amd.warm_start import LearnedWarmStart
    mesh, mu_t = physical_block(5, seed=seed)
    kw = dict(kw)
    if kw.pop("volume", False):
        nu = 0.45
        kw["volume_lambda"] = (mesh.E * nu / ((1 + nu) * (1 - 2 * nu))).astype(np.float32)
    args = solver_kwargs(mesh, mu_per_tet=mu_t, **kw)
    if start_deformed:
        args["v_p"] = deformed(mesh, seed=seed, amp=0.02, twist=0.05)
    solvers = []
    for native in (False, True):
        extra = {}
        if kw.get("warm_start") == "learned":
            n_cons = mesh.n_tets * (2 if "volume_lambda" in kw else 1)
            model = LearnedWarmStart.for_mesh(mesh.tets, mesh.n_tets, seed=3,
                                              per_tet=n_cons // mesh.n_tets)
            rng = np.random.default_rng(7)     # non-zero decoder: the prediction differs
            p = model.params.cpu().numpy()
            model.load_state_dict({"params": (p + 0.05 * rng.standard_normal(p.shape))
                                   .astype(np.float32)})
            extra = dict(warm_start_model=model)
        s = MGPBDSolverGPU(**args, **extra, native_substep=native, verbose=False)
        s.init()
        solvers.append(s)
    return mesh, solvers


def _same(a, b):
    return torch.equal(a.view(torch.int32), b.view(torch.int32))


@pytest.mark.parametrize("case", list(CASES))
def test_native_substep_is_bit_identical(case):
    kw, start_deformed, drag = CASES[case]
    mesh, (py, nat) = _pair(kw, start_deformed, seed=61)
    rng = np.random.default_rng(5)
    free = np.nonzero(mesh.invm > 0)[0]
    predicted = 0
    for k in range(6):
        if drag and k % 3 == 0:     # a new drag target every third substep
            v = int(rng.choice(free))
            target = mesh.verts[v] + rng.uniform(-0.05, 0.05, 3)
            for s in (py, nat):
                s.set_drag(v, target, k=1e3)
        print(f"[native] {case} substep {k}", flush=True)
        py.step(0)
        nat.step(0)
        assert nat._native is not None and py._native is None
        for name in ("v_p", "v_v", "lambdaf", "_lam_prev"):
            assert _same(getattr(py, name), getattr(nat, name)), (k, name)
        assert py.last_stats == nat.last_stats, (k, py.last_stats, nat.last_stats)
        assert py.last_solve_info == nat.last_solve_info, k
        assert py.last_predicted == nat.last_predicted
        predicted += nat.last_predicted
    print(f"[native] {case}: 6 substeps bit-identical; last {nat.last_stats}", flush=True)
    if kw.get("warm_start") == "learned":
        assert predicted == 6
        assert (nat._native[0] is not None) and ((py._side is None) ==
                                                 (not kw.get("overlap_inference", True)))


def test_lag_rescue_is_bit_identical_and_verified():
    """The lagged-hierarchy rescue (levels 1.. rebuilt only at a substep's first Newton
    iteration; a lagged solve past 2 iterations rebuilds them and resumes) on a stiff-contrast
    block under drag: it fires, the native and Python loops stay bit-identical, every solve is
    verified at 1e-6."""
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    mesh, mu_t = physical_block(6, seed=63, E_range=(1e4, 1e8))
    args = solver_kwargs(mesh, mu_per_tet=mu_t, **dict(MIXED, precision="fp16",
                                                       level0="factored",
                                                       coarse_refresh="substep"))
    args["v_p"] = deformed(mesh, seed=63, amp=0.02, twist=0.05)
    py, nat = (MGPBDSolverGPU(**args, native_substep=native, verbose=False)
               for native in (False, True))
    for s in (py, nat):
        s.init()
        s._gpu_solver.ctx("mgpbd_pcg_set_lag_rescue", 2)
    rng = np.random.default_rng(6)
    free = np.nonzero(mesh.invm > 0)[0]
    for k in range(4):
        v = int(rng.choice(free))
        target = mesh.verts[v] + rng.uniform(-0.05, 0.05, 3)
        for s in (py, nat):
            s.set_drag(v, target, k=1e3)
        py.step(0)
        nat.step(0)
        for name in ("v_p", "v_v", "lambdaf"):
            assert _same(getattr(py, name), getattr(nat, name)), (k, name)
        assert py.last_stats == nat.last_stats, k
        assert all(d["reason"] != 0 for d in nat.last_solve_info), k
    fired = nat._gpu_solver.lag_rescues()
    print(f"[lag rescue] 4 substeps bit-identical; lag rescues {fired} / "
          f"{py._gpu_solver.lag_rescues()}; last {nat.last_stats}", flush=True)
    assert fired > 0 and fired == py._gpu_solver.lag_rescues()


def test_native_substep_falls_back_while_hooks_are_attached():
    """With a SolverBenchmark attached the Python loop runs (its hooks are Python)."""
    from controllable_pbd_3d_amd.benchmark import SolverBenchmark
    mesh, (py, nat) = _pair({}, True, seed=62)
    nat.benchmark = SolverBenchmark("native", "box5", {})
    nat.step(0)
    assert nat._native is None
    assert nat.benchmark.summary()["n_frames"] == 1
