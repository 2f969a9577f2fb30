"""Per-frame, per-Newton-iteration timing and convergence record of a solve.

Wire format (the JSON document of the reference's ``SolverBenchmark``, utils/bench.py:14-97,
which its plotting reads): ``solver_name``, ``mesh_name``, ``params`` and ``frames``; a frame
holds ``frame_id``, ``t_start``, ``t_end`` (seconds) and ``iterations``, each with
``iteration``, ``residual``, ``cg_iters``, ``omega`` and ``t`` (seconds, same clock).

The hooks ``begin_frame`` / ``record_iteration`` / ``end_frame`` are the ones
``MGPBDSolverGPU`` calls when constructed with ``benchmark=``. Timing here does not stall the
solve: on a GPU every hook records a HIP event on the current stream and returns; the events
are turned into times only when the record is read (``frames``, ``save``, ``summary``), by
waiting for the last event of each pending frame. A frame's times are then the device
timeline between the hooks: the begin event anchors the frame at the host clock reading taken
when ``begin_frame`` was called, and every later mark adds the device time elapsed since that
event. Without a GPU (CPU solvers such as the oracle) the marks are host clock readings.
"""

from __future__ import annotations

import dataclasses
import json
import time


def _device_clock():
    """A factory of HIP timing events on the current device, or None without a GPU."""
    try:
        import torch
    except ImportError:
        return None
    if not (torch.cuda.is_available() and torch.cuda.is_initialized()):
        return None

    def mark():
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(torch.cuda.current_stream())
        return ev
    return mark


@dataclasses.dataclass
class _PendingFrame:
    frame_id: int
    host_t0: float                   # host clock at begin_frame
    start: object                    # event or host time of the begin mark
    marks: list                      # one mark per recorded iteration
    stats: list                      # (iteration, residual, cg_iters, omega)
    end: object = None

    def resolve(self) -> dict:
        if hasattr(self.start, "elapsed_time"):
            self.end.synchronize()   # waits for this frame's last event only

            def at(m):
                return self.host_t0 + self.start.elapsed_time(m) * 1e-3
        else:
            def at(m):
                return self.host_t0 + (m - self.start)
        its = [{"iteration": int(i), "residual": float(r), "cg_iters": int(c),
                "omega": float(w), "t": at(m)}
               for (i, r, c, w), m in zip(self.stats, self.marks)]
        return {"frame_id": self.frame_id, "iterations": its, "t_start": self.host_t0,
                "t_end": at(self.end)}


class SolverBenchmark:
    """Timing/convergence record with the reference's hooks and JSON document."""

    def __init__(self, solver_name: str, mesh_name: str, params: dict | None = None):
        self.solver_name = solver_name
        self.mesh_name = mesh_name
        self.params = dict(params or {})
        self._done = []              # resolved frame dicts
        self._pending = []           # finished frames whose events are not read yet
        self._open = None            # the frame between begin_frame and end_frame
        self._mark = None

    # -- hooks ------------------------------------------------------------------------
    def _now(self):
        if self._mark is None:
            self._mark = _device_clock() or time.perf_counter
        return self._mark()

    def begin_frame(self, frame_id: int):
        self._mark = None            # pick the clock per frame (a GPU may appear later)
        host_t0 = time.perf_counter()
        self._open = _PendingFrame(int(frame_id), host_t0, self._now(), [], [])

    def record_iteration(self, iteration: int, residual: float, cg_iters: int = 0,
                         omega: float = 1.0):
        fr = self._open
        if fr is None:
            return
        fr.marks.append(self._now())
        fr.stats.append((iteration, residual, cg_iters, omega))

    def end_frame(self):
        fr = self._open
        if fr is None:
            return
        fr.end = self._now()
        self._pending.append(fr)
        self._open = None

    # -- reading ----------------------------------------------------------------------
    @property
    def frames(self) -> list:
        if self._pending:
            self._done.extend(fr.resolve() for fr in self._pending)
            self._pending.clear()
        return self._done

    @frames.setter
    def frames(self, value):
        self._pending.clear()
        self._done = list(value)

    def to_dict(self) -> dict:
        return {"solver_name": self.solver_name, "mesh_name": self.mesh_name,
                "params": self.params, "frames": self.frames}

    def save(self, path: str):
        with open(path, "w") as fh:
            json.dump(self.to_dict(), fh, indent=2)

    @classmethod
    def load(cls, path: str) -> "SolverBenchmark":
        with open(path) as fh:
            doc = json.load(fh)
        rec = cls(doc["solver_name"], doc["mesh_name"], doc.get("params"))
        rec.frames = doc["frames"]
        return rec

    def summary(self) -> dict:
        """Means over the frames: frame time (frames with both ends recorded), Newton
        iterations per frame, last residual of each frame; and the summed frame time.
        Keys as utils/bench.py:76-97 returns them; empty when no frame was recorded."""
        frames = self.frames
        if not frames:
            return {}

        def mean(v):
            return sum(v) / len(v) if v else 0.0

        spans = [f["t_end"] - f["t_start"] for f in frames
                 if f.get("t_start") is not None and f.get("t_end") is not None]
        finals = [f["iterations"][-1]["residual"] for f in frames if f["iterations"]]
        return {"solver_name": self.solver_name, "mesh_name": self.mesh_name,
                "n_frames": len(frames),
                "avg_frame_time": float(mean(spans)),
                "avg_iterations": float(mean([len(f["iterations"]) for f in frames])),
                "avg_final_residual": float(mean(finals)),
                "total_time": float(sum(spans))}
