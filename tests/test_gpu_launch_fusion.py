"""Launch fusion of the reduced-precision V-cycle (mgpbd_mg_set_launch_fusion): prolongation from
the dense coarsest level inside its product, against the separate prolongation launch of the
reference's V-cycle (cons/mgpcg_gpu.py:251-253), and level 0's scaled rhs and first sweep in the
PCG update.
Two solvers built alike, one with the fusion switched off, stepped side by side: positions,
lambda and the per-iteration records must agree bit for bit after every substep."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import deformed, physical_block, solver_kwargs   # noqa: E402

EXT = dict(hierarchy="extended", galerkin="structural", max_coarse=16, relative_tol=True,
           tol_cg=1e-6, hold_pinned=True, verify_residual=1, fp64_iterate=True)

CASES = {
    # the benchmarked configuration: factored level 0, Chebyshev, dense coarsest (the dense
    # product pushes onto the level above it)
    "fp16_factored_chebyshev_dense": dict(EXT, precision="fp16", level0="factored",
                                          coarse_solver="cholesky", smoother="chebyshev"),
    # assembled level 0, spectral Jacobi, Jacobi sweeps on the coarsest level (only the PCG
    # update's fusion applies)
    "bf16_assembled_spectral_jacobi_coarsest": dict(EXT, precision="bf16", smoother="spectral",
                                                    coarse_solver="jacobi", coarse_sweeps=20),
    # W-cycle: the coarsest level pushes, the doubled levels prolong separately
    "fp16_wcycle": dict(EXT, precision="fp16", level0="factored", coarse_solver="cholesky",
                        smoother="spectral", cycle="W"),
}


def _same(a, b):
    return torch.equal(a.view(torch.int32), b.view(torch.int32))


@pytest.mark.parametrize("case", list(CASES))
def test_launch_fusion_is_bit_identical(case):
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    mesh, mu_t = physical_block(8, seed=71)
    args = solver_kwargs(mesh, mu_per_tet=mu_t, **CASES[case])
    args["v_p"] = deformed(mesh, seed=71, amp=0.02, twist=0.05)
    fused, plain = (MGPBDSolverGPU(**args, verbose=False) for _ in range(2))
    for s in (fused, plain):
        s.init()
    plain._gpu_solver.ctx("mgpbd_mg_set_launch_fusion", 0)
    sizes = fused._gpu_solver._level_sizes
    print(f"[launch fusion] {case}: levels {sizes}", flush=True)
    assert len(sizes) >= 4
    for k in range(4):
        fused.step(0)
        plain.step(0)
        for name in ("v_p", "v_v", "lambdaf"):
            assert _same(getattr(fused, name), getattr(plain, name)), (k, name)
        assert fused.last_stats == plain.last_stats, (k, fused.last_stats, plain.last_stats)
        assert fused.last_solve_info == plain.last_solve_info, k
    print(f"[launch fusion] {case}: 4 substeps bit-identical; last {fused.last_stats}",
          flush=True)
    assert all(d["reason"] != 0 for d in fused.last_solve_info)
