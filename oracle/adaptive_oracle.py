"""ORACLE (test infrastructure only) -- numpy restatement of the adaptive zone and the
adaptive wrapper's compliance mask (cons/adaptive_zone.py, cons/mgpbd_adaptive.py). Imported
by ``tests/`` only, as the checker; the product never calls it.

* ``proximity``   cons/adaptive_zone.py:35-48  FP32 ``(dx dx + dy dy) + dz dz < r * r``
* ``strain``      cons/adaptive_zone.py:50-64  FP32 ``|x - x_old| / dt > threshold``
* ``expand``      cons/adaptive_zone.py:66-80  the reference's loop over active tets, marking
                                             their face neighbours, over the dict-of-sets
                                             adjacency (oracle.structure)
* ``mask_alpha``  cons/mgpbd_adaptive.py:64-79
* ``adaptive_step`` cons/mgpbd_adaptive.py:36-62 around an ``OracleMGPBD``

Pinning (tests/test_adaptive.py): proximity and strain against FP64 evaluations away from
ties; ``expand`` against breadth-first face-graph distance (<= n rings); the mask against
its definition.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def proximity(t_i, v_p, center, radius):
    x = np.asarray(v_p, F32).reshape(-1, 3)
    t = np.asarray(t_i).reshape(-1, 4)
    c = np.asarray(center, F32)
    r2 = F32(radius) * F32(radius)
    p = x[t]                                     # (n, 4, 3)
    d = p - c
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return (d2 < r2).any(axis=1).astype(np.int32)


def strain(t_i, v_p, v_p_old, threshold, dt):
    x = np.asarray(v_p, F32).reshape(-1, 3)
    xo = np.asarray(v_p_old, F32).reshape(-1, 3)
    t = np.asarray(t_i).reshape(-1, 4)
    d = x[t] - xo[t]
    n2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    vel = np.sqrt(n2) / F32(dt)
    vmax = np.maximum(vel.max(axis=1), F32(0))
    return (vmax > F32(threshold)).astype(np.int32)


def expand(adj: dict, active, n_rings):
    active = np.asarray(active, np.int32).copy()
    for _ in range(n_rings):
        new = active.copy()
        for t in range(len(active)):
            if active[t] == 1:
                for nb in adj[t]:
                    new[nb] = 1
        active = new
    return active


def mask_alpha(active, alpha, passive_value=1e20):
    saved = np.asarray(alpha, F32).copy()
    out = saved.copy()
    out[np.asarray(active) == 0] = F32(passive_value)
    return out, saved


def adaptive_step(o, adj, active, frame=0, tool_pos=None, tool_radius=0.5):
    """One adaptive substep of an OracleMGPBD ``o``; returns the zone used."""
    o.make_prediction()
    if tool_pos is not None:
        active = proximity(o.t_i, o.v_p, tool_pos, tool_radius)
        active = expand(adj, active, 1)
    masked, saved = mask_alpha(active, o.alpha_tilde)
    o.alpha_tilde[:] = masked
    o.solve(frame)
    o.alpha_tilde[:] = saved
    o.update_vel()
    return active
