"""MI355X-native MGPBD global XPBD solve (drop-in for the reference's MGPBDSolverGPU).

Public API mirrors the reference's cons/ modules:
  MGPBDSolverGPU      (cons/mgpbd_gpu.py)
  GPUMultigridSolver  (cons/mgpcg_gpu.py)
  GPUSparseMatrix, gpu_* BLAS-1 (cons/sparse_gpu.py)
  tet_aggregation.*   (cons/tet_aggregation.py)
  AdaptiveZone, MGPBDAdaptiveSolver (cons/adaptive_zone.py, cons/mgpbd_adaptive.py)
  SolverBenchmark     (utils/bench.py)
"""

from ._lib import MGPBDError, load as load_library  # noqa: F401


def __getattr__(name):
    # torch-dependent modules load on first use so that host-only tools import cheaply
    if name == "MGPBDSolverGPU":
        from .mgpbd_gpu import MGPBDSolverGPU
        return MGPBDSolverGPU
    if name == "GPUMultigridSolver":
        from .mgpcg_gpu import GPUMultigridSolver
        return GPUMultigridSolver
    if name == "GPUSparseMatrix":
        from .sparse_gpu import GPUSparseMatrix
        return GPUSparseMatrix
    if name in ("AdaptiveZone", "MGPBDAdaptiveSolver"):
        from . import adaptive
        return getattr(adaptive, name)
    if name == "SolverBenchmark":
        from .benchmark import SolverBenchmark
        return SolverBenchmark
    raise AttributeError(name)
