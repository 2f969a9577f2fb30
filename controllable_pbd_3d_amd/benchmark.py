"""Per-frame, per-Newton-iteration timing and convergence record (utils/bench.py:14-97).

Same hooks (``begin_frame``, ``record_iteration``, ``end_frame``) that ``MGPBDSolverGPU``
calls when constructed with ``benchmark=``, and the same JSON document (``solver_name``,
``mesh_name``, ``params``, ``frames`` with ``frame_id``, ``iterations`` [``iteration``,
``residual``, ``cg_iters``, ``omega``, ``t``], ``t_start``, ``t_end``), so files written here
load in the reference's plotting and the reference's load here. Every hook waits for the GPU
first (the reference calls ``ti.sync()``), so the times bracket finished device work.
"""

from __future__ import annotations

import json
import time

import numpy as np


def _sync():
    try:
        import torch
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.synchronize()
    except ImportError:
        pass


class SolverBenchmark:
    def __init__(self, solver_name: str, mesh_name: str, params: dict | None = None):
        self.solver_name = solver_name
        self.mesh_name = mesh_name
        self.params = params or {}
        self.frames = []
        self._current_frame = None

    def begin_frame(self, frame_id: int):
        _sync()
        self._current_frame = {"frame_id": frame_id, "iterations": [],
                               "t_start": time.perf_counter(), "t_end": None}

    def record_iteration(self, iteration: int, residual: float, cg_iters: int = 0,
                         omega: float = 1.0):
        if self._current_frame is None:
            return
        _sync()
        self._current_frame["iterations"].append({
            "iteration": iteration, "residual": float(residual), "cg_iters": cg_iters,
            "omega": float(omega), "t": time.perf_counter()})

    def end_frame(self):
        if self._current_frame is None:
            return
        _sync()
        self._current_frame["t_end"] = time.perf_counter()
        self.frames.append(self._current_frame)
        self._current_frame = None

    def save(self, path: str):
        with open(path, "w") as f:
            json.dump({"solver_name": self.solver_name, "mesh_name": self.mesh_name,
                       "params": self.params, "frames": self.frames}, f, indent=2)

    @staticmethod
    def load(path: str) -> "SolverBenchmark":
        with open(path) as f:
            data = json.load(f)
        b = SolverBenchmark(data["solver_name"], data["mesh_name"], data.get("params", {}))
        b.frames = data["frames"]
        return b

    def summary(self) -> dict:
        """Means over the recorded frames (utils/bench.py:76-97): frame time, Newton
        iterations, final residual; total time."""
        if not self.frames:
            return {}
        times, iters, finals = [], [], []
        for fr in self.frames:
            if fr["t_end"] is not None and fr["t_start"] is not None:
                times.append(fr["t_end"] - fr["t_start"])
            its = fr["iterations"]
            iters.append(len(its))
            if its:
                finals.append(its[-1]["residual"])
        return {"solver_name": self.solver_name, "mesh_name": self.mesh_name,
                "n_frames": len(self.frames),
                "avg_frame_time": float(np.mean(times)) if times else 0.0,
                "avg_iterations": float(np.mean(iters)) if iters else 0.0,
                "avg_final_residual": float(np.mean(finals)) if finals else 0.0,
                "total_time": float(np.sum(times)) if times else 0.0}
