"""Shared inputs for the parity tests: small jittered blocks and deformed starting states."""

from __future__ import annotations

import numpy as np

from controllable_pbd_3d_amd.mesh import box_mesh


def deformed(mesh, seed=0, amp=0.15, twist=0.3):
    """Rest positions twisted about the vertical axis and stretched, plus noise of
    ``amp * h``, so that every tet starts strained (C > 1e-6) and A has off-diagonals."""
    rng = np.random.default_rng(1000 + seed)
    x = mesh.verts.astype(np.float64).copy()
    c = x.mean(axis=0)
    y = (x[:, 1] - x[:, 1].min()) / max(np.ptp(x[:, 1]), 1e-12)
    ang = twist * (y - 0.5)
    dx, dz = x[:, 0] - c[0], x[:, 2] - c[2]
    x[:, 0] = c[0] + np.cos(ang) * dx - np.sin(ang) * dz
    x[:, 2] = c[2] + np.sin(ang) * dx + np.cos(ang) * dz
    x[:, 1] = c[1] + 1.08 * (x[:, 1] - c[1])
    x += rng.uniform(-amp * mesh.h, amp * mesh.h, size=x.shape)
    return x.astype(np.float32)


def solver_kwargs(mesh, **over):
    kw = dict(v_p=mesh.verts.copy(), v_p_ref=mesh.verts.copy(), v_invm=mesh.invm.copy(),
              t_i=mesh.tets.copy(), t_m=mesh.tet_mass.copy(), gravity=(0.0, -9.8, 0.0),
              dt=1.0 / 60 / 5, mu=1e6, damp=1.0, maxiter=20, atol=1e-4, rtol=1e-2,
              maxiter_cg=100, tol_cg=1e-5, n_smooth=2, omega_jacobi=2.0 / 3.0)
    kw.update(over)
    return kw


def small_block(n=5, seed=0, h=0.05, **kw):
    return box_mesh(n, h, jitter=0.05, seed=seed, **kw)


def physical_block(n=5, seed=0, h=0.05, E_range=(1e4, 1e6), nu=0.45, rho=1000.0):
    """Block with the workloads' material model: rho-scaled masses, per-tet E log-uniform,
    mu_t = E_t / (2 (1 + nu)) (Lame mu, the convention of cons/vbd.py:44-47)."""
    mesh = box_mesh(n, h, jitter=0.05, seed=seed, rho=rho, E_range=E_range)
    mu_t = (mesh.E / (2.0 * (1.0 + nu))).astype(np.float32)
    return mesh, mu_t
