"""The benchmark workloads of the MGPBD solve: meshes, materials and loading.

* ``c0``: 8^3 Kuhn block (3,072 tets), h 0.05 m, uniform E 1e5 Pa;
* ``c1``: 56^3 block (1,053,696 tets), h 0.02 m, per-tet E log-uniform in [1e4, 1e6] Pa;
* ``c2``: the ``c1`` block with E log-uniform in [1e4, 1e8] Pa;
* ``c3``: 32^3 block (196,608 tets), h 0.03 m, E in [1e4, 1e6] Pa, spring drag;
* ``c4``: jittered grid (h 0.01 m) clipped to an ellipsoid with radii (1.0, 0.6, 0.4) m
  (~6.0M tets), a stiff sphere of radius 0.2 m (E 1e8) in an E 1e5 matrix, lowest 5 % of the
  vertices held, spring drag.

Every workload: interior vertices jittered by U(+-0.05 h) from the seed, rho 1000 kg/m^3,
nu 0.45, gravity -9.8 m/s^2, top face held (boxes), dt 1/60 s with 5 substeps. Lame
parameters per tet from E and nu (the convention of cons/vbd.py:44-47): mu = E / (2 (1 + nu))
for the ARAP constraint, lambda = E nu / ((1 + nu)(1 - 2 nu)) for the volume constraint.
Spring drag: every frame a seeded random free vertex is pulled by a k = 1e3 N/m spring toward
a target uniform in a 0.5 m cube around its rest position.
"""

from __future__ import annotations

import numpy as np

NU = 0.45
RHO = 1000.0
SUBSTEPS_PER_FRAME = 5
DT = 1.0 / 60.0 / SUBSTEPS_PER_FRAME
DRAG_K, DRAG_BOX = 1e3, 0.5

# kind box: n^3 cubes of edge h; E (lo, hi) log-uniform per tet, None = uniform 1e5.
# max_levels: levels of the extended hierarchy; maxiter_cg: PCG cap where some systems need
# more than the reference's 100 iterations to reach 1e-6; smoothing: Chebyshev steps per
# phase on level 0 and on the coarse levels and the interval ratio (measured per workload,
# default 5 / 1 on [1.2 lambda / 30, 1.2 lambda]); coarse_refresh: when levels 1.. are rebuilt
# (measured per workload: the inclusion under drag moves level 0 so far at every Newton
# iteration that rebuilding each time beats the adaptive policy).
CONFIGS = {
    "c0": dict(kind="box", n=8, h=0.05, E=None, max_levels=3, label="box8_3072tets"),
    "c1": dict(kind="box", n=56, h=0.02, E=(1e4, 1e6), max_levels=5, label="box56_1053696tets",
               smoothing=((13, 1), 40.0)),
    "c2": dict(kind="box", n=56, h=0.02, E=(1e4, 1e8), max_levels=5,
               label="box56_1053696tets_E1e4-1e8", smoothing=((8, 1), 40.0)),
    "c3": dict(kind="box", n=32, h=0.03, E=(1e4, 1e6), max_levels=4, drag=True,
               label="box32_196608tets_drag", smoothing=((8, 1), 40.0)),
    "c4": dict(kind="ellipsoid", h=0.01, E=(1e5, 1e8), max_levels=6, drag=True,
               maxiter_cg=1000, coarse_refresh="newton",
               label="ellipsoid_h0.01_inclusion_drag"),
}


def lame(E, nu: float = NU):
    """(mu, lambda) per tet, float32 (cons/vbd.py:44-47)."""
    E = np.asarray(E, np.float64)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu.astype(np.float32), lam.astype(np.float32)


def build_workload(cfg, seed: int = 0):
    """(mesh, mu per tet, lambda per tet) of a workload (a CONFIGS entry or its name)."""
    from .mesh import box_mesh, ellipsoid_mesh
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    if cfg["kind"] == "ellipsoid":
        mesh = ellipsoid_mesh(cfg["h"], jitter=0.05, seed=seed, rho=RHO,
                              E_matrix=cfg["E"][0], E_inclusion=cfg["E"][1])
    else:
        mesh = box_mesh(cfg["n"], cfg["h"], jitter=0.05, seed=seed, rho=RHO,
                        E_range=cfg["E"] if cfg["E"] else (1e5, 1e5))
    mu_t, lam_t = lame(mesh.E)
    return mesh, mu_t, lam_t


def solver_args(mesh, mu_t, lam_t=None, volume: bool = False, **over):
    """Keyword arguments of MGPBDSolverGPU for a workload mesh: the reference's Newton defaults
    (maxiter 20, atol 1e-4, rtol 1e-2, line search), the workloads' PCG (relative 1e-6, 100
    iterations, 2 + 2 sweeps, omega 2/3), held pins; the volume constraint when ``volume``."""
    kw = dict(v_p=mesh.verts.copy(), v_p_ref=mesh.verts.copy(), v_invm=mesh.invm.copy(),
              t_i=mesh.tets.copy(), t_m=mesh.tet_mass.copy(), gravity=(0.0, -9.8, 0.0),
              dt=DT, maxiter=20, atol=1e-4, rtol=1e-2, maxiter_cg=100, tol_cg=1e-6,
              relative_tol=True, n_smooth=2, omega_jacobi=2.0 / 3.0, use_line_search=True,
              mu_per_tet=mu_t, hold_pinned=True)
    if volume:
        kw["volume_lambda"] = lam_t
    kw.update(over)
    return kw


class Loading:
    """Per-substep loading of a workload: nothing, or the drag schedule of ``seed`` (one target
    per frame of SUBSTEPS_PER_FRAME substeps)."""

    def __init__(self, cfg, mesh, seed: int, substeps: int):
        from .mesh import drag_schedule
        if isinstance(cfg, str):
            cfg = CONFIGS[cfg]
        self.frames = None
        if cfg.get("drag"):
            self.frames = drag_schedule(mesh, seed, substeps // SUBSTEPS_PER_FRAME + 1, DRAG_BOX)
        self.k = 0

    def step(self, s):
        if self.frames is not None and self.k % SUBSTEPS_PER_FRAME == 0:
            v, target = self.frames[self.k // SUBSTEPS_PER_FRAME]
            s.set_drag(v, target, DRAG_K)
        s.step(self.k // SUBSTEPS_PER_FRAME)
        self.k += 1
