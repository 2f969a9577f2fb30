"""Error paths of the substep, forced from the host: a NaN written into the solver's state
must come back as MGPBD_ENONFINITE instead of running every Newton / PCG iteration to its cap
(the reference compares a NaN residual and simply continues, cons/mgpbd_gpu.py:410-417 and
cons/mgpcg_gpu.py:305-306). Every kernel still runs to completion (NaN arithmetic only)."""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import physical_block, solver_kwargs   # noqa: E402


def _solver(**kw):
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    mesh, mu_t = physical_block(4, seed=31)
    s = MGPBDSolverGPU(**solver_kwargs(mesh, mu_per_tet=mu_t, hold_pinned=True, **kw),
                       verbose=False)
    s.init()
    return mesh, s


@pytest.mark.parametrize("native", [False, True])
@pytest.mark.parametrize("mode", ["reference", "extended_fp16"])
def test_nan_position_reports_nonfinite(mode, native):
    """Python Newton loop and the library's mgpbd_substep alike."""
    from controllable_pbd_3d_amd._lib import MGPBD_ENONFINITE, MGPBDError
    kw = {} if mode == "reference" else dict(
        hierarchy="extended", galerkin="structural", relative_tol=True, tol_cg=1e-6,
        max_coarse=64, coarse_solver="cholesky", smoother="chebyshev", verify_residual=1,
        fp64_iterate=True, precision="fp16", level0="factored")
    mesh, s = _solver(native_substep=native, **kw)
    s.step(0)                                   # a healthy substep first
    free = int(np.nonzero(mesh.invm > 0)[0][0])
    s.v_p[free, 1] = float("nan")               # host write into the device state
    print(f"[errors] {mode} native={native}: NaN at vertex {free}", flush=True)
    with pytest.raises(MGPBDError) as e:
        s.step(1)
    assert "non-finite" in str(e.value)
    assert str(MGPBD_ENONFINITE) in str(e.value)


def test_pcg_with_nan_rhs_returns_enonfinite():
    from controllable_pbd_3d_amd._lib import MGPBD_ENONFINITE, MGPBDError
    mesh, s = _solver()
    b = torch.ones(s.n_tets, dtype=torch.float32, device=s.device)
    b[s.n_tets // 2] = float("nan")
    x = torch.zeros_like(b)
    with pytest.raises(MGPBDError) as e:
        s._gpu_solver.solve(b, x)
    assert e.value.status == MGPBD_ENONFINITE
    assert "non-finite" in str(e.value)
    # the context stays usable: a finite solve afterwards converges
    b[s.n_tets // 2] = 1.0
    it = s._gpu_solver.solve(b, x)
    assert it > 0 and math.isfinite(float(x.abs().max()))
