"""The factored level 0's layouts: for the vertex passes (mgpbd_mg_set_level0_offsets) the tet
of each slot from the incidence list, from a 16-bit offset beside its K / gradient, and the
uniform layout (every vertex slice as tall as the busiest vertex, slot positions computed,
unused slots zero); for the tet passes (mgpbd_mg_set_level0_packed_rows) the row vertices from
the int4 list or from one packed 64-bit word. Same tets, same summation order, zeros added
last: positions bit-identical across all of them. The host-side base, layout and packing are
checked on the CPU in tests/test_slot_offsets.py."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import deformed, physical_block, solver_kwargs   # noqa: E402


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_slot_layouts_are_bit_identical(monkeypatch, precision):
    from controllable_pbd_3d_amd import MGPBDSolverGPU, mgpcg_gpu
    from controllable_pbd_3d_amd.mgpcg_gpu import GPUMultigridSolver
    mesh, mu_t = physical_block(6, seed=71)
    kw = solver_kwargs(mesh, mu_per_tet=mu_t, hold_pinned=True, relative_tol=True, tol_cg=1e-6,
                       hierarchy="extended", galerkin="structural", max_coarse=64,
                       coarse_solver="cholesky", smoother="chebyshev", verify_residual=1,
                       fp64_iterate=True, rtol=1e-6, atol=1e-9, maxiter=30,
                       precision=precision, level0="factored")
    kw["v_p"] = deformed(mesh, seed=71, amp=0.02, twist=0.05)
    runs = []
    uniform_rows = mgpcg_gpu.uniform_rows
    for layout in ("incidence", "offsets", "uniform", "uniform_packed"):
        monkeypatch.setattr(GPUMultigridSolver, "slot_offsets", layout != "incidence")
        monkeypatch.setattr(GPUMultigridSolver, "pack_rows", layout == "uniform_packed")
        monkeypatch.setattr(mgpcg_gpu, "uniform_rows",
                            uniform_rows if layout.startswith("uniform") else (lambda vptr: 0))
        s = MGPBDSolverGPU(**kw, verbose=False)
        s.init()
        f = s._gpu_solver._fac
        assert ("tbase" in f) == (layout != "incidence")
        assert (f.get("rows_uniform", 0) > 0) == layout.startswith("uniform")
        assert ("rows_packed" in f) == (layout == "uniform_packed")
        for _ in range(3):
            s.step(0)
        runs.append((s.v_p.cpu().numpy(), s.last_stats))
    print(f"[layouts] {precision}: {runs[-1][1]}", flush=True)
    for v, st in runs[1:]:
        assert np.array_equal(runs[0][0].view(np.int32), v.view(np.int32))
        assert runs[0][1] == st
