"""Spatial adaptivity around the MGPBD solve: the active/passive tet zone and the wrapper that
masks passive tets out of a substep (cons/adaptive_zone.py, cons/mgpbd_adaptive.py).

The zone is one int32 flag per tet held on the GPU (``tet_active``); the proximity, strain,
ring-expansion and compliance-mask steps are libmgpbd kernels. The wrapper inflates alpha~
of passive tets to 1e20 for the solve (their constraints stop pulling) and restores the
original values before the velocity update, as the reference does.
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import dptr
from .runtime import get_context, to_device
from .warm_start import neighbour_table


class AdaptiveZone:
    """Active/passive flags per tet (cons/adaptive_zone.py:14-94).

    ``t_i`` (4 n_tets int32, flat) and ``v_p`` (n_verts x 3 float32) are the solver's arrays;
    device tensors are used in place (the zone reads the solver's current positions), host
    arrays are copied once.
    """

    def __init__(self, n_tets: int, t_i, v_p, device=None):
        self.ctx = get_context(device)
        dev = self.ctx.device
        self.n_tets = int(n_tets)
        self.t_i = to_device(t_i, torch.int32, dev, (-1,))
        self.v_p = to_device(v_p, torch.float32, dev, (-1, 3))
        self.n_verts = int(self.v_p.shape[0])
        if self.t_i.numel() != 4 * self.n_tets:
            raise ValueError("t_i must hold 4 vertex ids per tet")
        # every tet active by default (cons/adaptive_zone.py:27-29)
        self.tet_active = torch.ones(self.n_tets, dtype=torch.int32, device=dev)
        self._scratch = torch.empty_like(self.tet_active)
        self._adj = None   # face neighbours, built on the first expansion

    def update_by_proximity(self, tool_x: float, tool_y: float, tool_z: float, radius: float):
        """Active = some vertex strictly within ``radius`` of the tool (FP32 distances)."""
        self.ctx("mgpbd_zone_proximity", self.n_tets, dptr(self.t_i), dptr(self.v_p),
                 float(tool_x), float(tool_y), float(tool_z), float(radius),
                 dptr(self.tet_active))

    def update_by_strain(self, v_p_old, threshold: float, dt: float):
        """Active = the fastest vertex moves faster than ``threshold`` (|x - x_old| / dt)."""
        old = to_device(v_p_old, torch.float32, self.ctx.device, (-1, 3))
        self.ctx("mgpbd_zone_strain", self.n_tets, dptr(self.t_i), dptr(self.v_p), dptr(old),
                 float(threshold), float(dt), dptr(self.tet_active))

    def expand_zone(self, n_rings: int):
        """Add ``n_rings`` rings of face-adjacent tets to the zone."""
        if self._adj is None:
            t_np = self.t_i.cpu().numpy()
            self._adj = torch.from_numpy(neighbour_table(t_np, self.n_tets)).to(self.ctx.device)
        self.ctx("mgpbd_zone_expand", self.n_tets, dptr(self._adj), dptr(self.tet_active),
                 dptr(self._scratch), int(n_rings))

    def get_active_mask(self) -> np.ndarray:
        return self.tet_active.cpu().numpy().astype(bool)

    def get_active_fraction(self) -> float:
        return float(self.get_active_mask().sum()) / max(self.n_tets, 1)


class MGPBDAdaptiveSolver:
    """Adds spatial adaptivity to an ``MGPBDSolverGPU`` (cons/mgpbd_adaptive.py:14-93):
    passive tets get alpha~ = 1e20 for the solve of each substep."""

    PASSIVE_ALPHA = 1e20

    def __init__(self, solver, zone: AdaptiveZone, passive_iters: int = 2):
        self.solver = solver
        self.zone = zone
        self.passive_iters = passive_iters   # kept for interface parity; unused there too
        self.n_tets = solver.n_tets
        self.n_verts = solver.n_verts
        self.v_p = solver.v_p
        self.v_v = solver.v_v
        self.v_p_old = solver.v_p_old
        self.damp = solver.damp
        self._original_alpha = torch.empty_like(solver.alpha_tilde)

    def step(self, frame: int = 0, tool_pos=None, tool_radius: float = 0.5):
        """Prediction, zone update around ``tool_pos`` (plus a one-ring buffer), masked solve,
        restore, velocity update (cons/mgpbd_adaptive.py:36-62)."""
        self.solver.make_prediction()
        if tool_pos is not None:
            self.zone.update_by_proximity(float(tool_pos[0]), float(tool_pos[1]),
                                          float(tool_pos[2]), float(tool_radius))
            self.zone.expand_zone(1)
        self._apply_active_mask()
        self.solver.solve(frame)
        self._restore_mask()
        self.solver.update_vel()

    def _apply_active_mask(self):
        s = self.solver
        per = getattr(s, "per_tet", 1)
        # with the volume constraint both rows of a tet follow the tet's flag
        flags = self.zone.tet_active if per == 1 else self.zone.tet_active.repeat_interleave(per)
        s.ctx("mgpbd_zone_mask_alpha", per * self.n_tets, dptr(flags),
              dptr(s.alpha_tilde), dptr(self._original_alpha), self.PASSIVE_ALPHA)
        if hasattr(s, "_alpha_tilde_np"):
            s._alpha_tilde_np = s.alpha_tilde.cpu().numpy()

    def _restore_mask(self):
        s = self.solver
        s.ctx("mgpbd_copy", s.alpha_tilde.numel(), dptr(self._original_alpha),
              dptr(s.alpha_tilde))
        if hasattr(s, "_alpha_tilde_np"):
            s._alpha_tilde_np = s.alpha_tilde.cpu().numpy()
