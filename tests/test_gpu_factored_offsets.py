"""The factored level 0's 16-bit slot offsets (mgpbd_mg_set_level0_offsets): the vertex passes
take each slot's tet from the offset word beside its K / gradient instead of the incidence
list. Same tets, same summation order: positions bit-identical to the incidence-list passes.
The host-side base (slot_offset_base) is checked on the CPU in tests/test_slot_offsets.py."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import deformed, physical_block, solver_kwargs   # noqa: E402


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_slot_offsets_are_bit_identical(monkeypatch, precision):
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    from controllable_pbd_3d_amd.mgpcg_gpu import GPUMultigridSolver
    mesh, mu_t = physical_block(6, seed=71)
    kw = solver_kwargs(mesh, mu_per_tet=mu_t, hold_pinned=True, relative_tol=True, tol_cg=1e-6,
                       hierarchy="extended", galerkin="structural", max_coarse=64,
                       coarse_solver="cholesky", smoother="chebyshev", verify_residual=1,
                       fp64_iterate=True, rtol=1e-6, atol=1e-9, maxiter=30,
                       precision=precision, level0="factored")
    kw["v_p"] = deformed(mesh, seed=71, amp=0.02, twist=0.05)
    runs = []
    for offsets in (False, True):
        monkeypatch.setattr(GPUMultigridSolver, "slot_offsets", offsets)
        s = MGPBDSolverGPU(**kw, verbose=False)
        s.init()
        assert ("tbase" in s._gpu_solver._fac) == offsets
        for _ in range(3):
            s.step(0)
        runs.append((s.v_p.cpu().numpy(), s.last_stats))
    print(f"[offsets] {precision}: {runs[1][1]}", flush=True)
    assert np.array_equal(runs[0][0].view(np.int32), runs[1][0].view(np.int32))
    assert runs[0][1] == runs[1][1]
