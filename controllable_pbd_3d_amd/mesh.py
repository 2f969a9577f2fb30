"""Tet-mesh inputs for the MGPBD solve: synthetic jittered blocks, Medit files, masses.

The solver consumes the same arrays the reference's ``TetData`` holds
(``data/tet_data.py:29-44``):

* ``verts``  float32 (n_verts, 3) rest positions (``v_p`` / ``v_p_ref``),
* ``tets``   int32   (4 * n_tets,) flat vertex ids (``t_i``),
* ``invm``   float32 (n_verts,) inverse vertex masses (``v_invm``), 0 = pinned,
* ``tet_mass`` float32 (n_tets,) tet volumes (``t_m``).

Mass convention (``utils/geom3d.py:48-59``): each tet adds ``rho * |det| / 6 / 4`` to each
of its four vertices, ``rho = 1`` in the reference; the synthetic workloads pass ``rho``.
"""

from __future__ import annotations

import dataclasses
import numpy as np


@dataclasses.dataclass
class TetMesh:
    verts: np.ndarray          # float32 (n_verts, 3)
    tets: np.ndarray           # int32 (4 * n_tets,)
    invm: np.ndarray           # float32 (n_verts,)
    tet_mass: np.ndarray       # float32 (n_tets,) rest volumes (reference ``t_m``)
    pinned: np.ndarray         # bool (n_verts,)
    h: float = 0.0
    E: np.ndarray | None = None  # float32 (n_tets,) Young's modulus per tet (extension)
    name: str = "mesh"

    @property
    def n_verts(self) -> int:
        return int(self.verts.shape[0])

    @property
    def n_tets(self) -> int:
        return int(self.tets.shape[0] // 4)


# Kuhn / Freudenthal split of the unit cube along its main diagonal (0,0,0)->(1,1,1):
# one tet per permutation of the axes, walking the cube edges in that axis order.
def _kuhn_tets():
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    out = []
    for p in perms:
        c = [0, 0, 0]
        path = [tuple(c)]
        for ax in p:
            c[ax] = 1
            path.append(tuple(c))
        a, b, cc, d = [np.array(v, dtype=np.float64) for v in path]
        det = np.linalg.det(np.stack([b - a, cc - a, d - a], axis=1))
        if det < 0:  # keep every tet positively oriented
            path[1], path[2] = path[2], path[1]
        out.append(path)
    return out


_KUHN = _kuhn_tets()


def vertex_mass(verts: np.ndarray, tets: np.ndarray, rho: float = 1.0):
    """Lumped vertex masses and tet volumes (``utils/geom3d.py:48-59``), FP64 accumulation."""
    t = tets.reshape(-1, 4).astype(np.int64)
    x = verts.astype(np.float64)
    x1, x2, x3, x4 = x[t[:, 0]], x[t[:, 1]], x[t[:, 2]], x[t[:, 3]]
    vol = np.abs(np.einsum("ij,ij->i", np.cross(x2 - x1, x3 - x1), x4 - x1)) / 6.0
    vm = np.zeros(x.shape[0], dtype=np.float64)
    for l in range(4):
        np.add.at(vm, t[:, l], rho * vol / 4.0)
    return vm, vol


def box_mesh(n: int, h: float, jitter: float = 0.05, seed: int = 0, rho: float = 1.0,
             E_range: tuple[float, float] | None = None, pin: str = "top",
             origin=(0.0, 0.0, 0.0), ny: int | None = None, nz: int | None = None) -> TetMesh:
    """Regular block of ``n x ny x nz`` cubes of edge ``h`` (ny = nz = n by default), 6 Kuhn
    tets per cube.

    * vertex id ``(ix * (ny+1) + iy) * (nz+1) + iz``; y is up (gravity along -y);
    * tet id ``6 * ((cx * ny + cy) * nz + cz) + k``;
    * interior vertices (not on the block boundary) are moved by U(-jitter*h, +jitter*h) per
      axis from ``np.random.default_rng(seed)``; the same generator then draws per-tet E
      log-uniform in ``E_range`` when given;
    * ``pin='top'`` pins the max-y face (``v_invm = 0``), as the reference demos pin ids.
    """
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    grid = np.stack([ix, iy, iz], axis=-1).reshape(-1, 3)
    verts = grid.astype(np.float64) * h + np.asarray(origin, dtype=np.float64)
    rng = np.random.default_rng(seed)
    interior = ((grid[:, 0] > 0) & (grid[:, 0] < nx) & (grid[:, 1] > 0) & (grid[:, 1] < ny)
                & (grid[:, 2] > 0) & (grid[:, 2] < nz))
    if jitter > 0:
        d = rng.uniform(-jitter * h, jitter * h, size=verts.shape)
        verts[interior] += d[interior]

    def vid(a, b, c):
        return (a * (ny + 1) + b) * (nz + 1) + c

    cx, cy, cz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cx, cy, cz = cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)
    tets = np.empty((cx.shape[0], 6, 4), dtype=np.int64)
    for k, path in enumerate(_KUHN):
        for l, (dx, dy, dz) in enumerate(path):
            tets[:, k, l] = vid(cx + dx, cy + dy, cz + dz)
    tets = tets.reshape(-1).astype(np.int32)

    verts32 = verts.astype(np.float32)
    vm, vol = vertex_mass(verts32, tets, rho)
    pinned = np.zeros(verts.shape[0], dtype=bool)
    if pin == "top":
        pinned = grid[:, 1] == ny
    elif pin != "none":
        raise ValueError(f"unknown pin mode {pin!r}")
    invm = (1.0 / vm).astype(np.float32)
    invm[pinned] = 0.0
    E = None
    if E_range is not None:
        lo, hi = np.log(E_range[0]), np.log(E_range[1])
        E = np.exp(rng.uniform(lo, hi, size=tets.shape[0] // 4)).astype(np.float32)
    return TetMesh(verts=verts32, tets=tets, invm=invm, tet_mass=vol.astype(np.float32),
                   pinned=pinned, h=h, E=E, name=f"box{nx}x{ny}x{nz}")


def ellipsoid_mesh(h: float, radii=(1.0, 0.6, 0.4), jitter: float = 0.05, seed: int = 0,
                   rho: float = 1.0, pin_fraction: float = 0.05,
                   inclusion_radius: float = 0.2, E_inclusion: float = 1e8,
                   E_matrix: float = 1e5) -> TetMesh:
    """Jittered Kuhn grid clipped to an ellipsoid (cubes whose unjittered centre lies inside
    are kept, so the connectivity does not depend on the seed), with a stiff spherical
    inclusion at the centre and the lowest ``pin_fraction`` of vertices by height pinned.
    Unused grid vertices are dropped and ids compacted."""
    rx, ry, rz = radii
    nx, ny, nz = [int(np.ceil(2 * r / h)) for r in radii]
    origin = (-nx * h / 2, -ny * h / 2, -nz * h / 2)
    full = box_mesh(nx, h, jitter=jitter, seed=seed, rho=rho, pin="none", origin=origin,
                    ny=ny, nz=nz)
    t = full.tets.reshape(-1, 4)
    cx, cy, cz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cube_c = (np.stack([cx, cy, cz], axis=-1).reshape(-1, 3) + 0.5) * h + np.asarray(origin)
    inside = ((cube_c[:, 0] / rx) ** 2 + (cube_c[:, 1] / ry) ** 2 + (cube_c[:, 2] / rz) ** 2) <= 1.0
    keep = np.repeat(inside, 6)
    t = t[keep]
    used = np.zeros(full.n_verts, dtype=bool)
    used[t.reshape(-1)] = True
    remap = -np.ones(full.n_verts, dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    verts = full.verts[used]
    tets = remap[t].reshape(-1).astype(np.int32)
    vm, vol = vertex_mass(verts, tets, rho)
    y = verts[:, 1]
    n_pin = max(1, int(np.floor(pin_fraction * verts.shape[0])))
    order = np.argsort(y, kind="stable")
    pinned = np.zeros(verts.shape[0], dtype=bool)
    pinned[order[:n_pin]] = True
    invm = (1.0 / vm).astype(np.float32)
    invm[pinned] = 0.0
    cen = verts.astype(np.float64)[tets.reshape(-1, 4)].mean(axis=1)
    E = np.where(np.linalg.norm(cen, axis=1) <= inclusion_radius, E_inclusion, E_matrix)
    return TetMesh(verts=verts, tets=tets, invm=invm, tet_mass=vol.astype(np.float32),
                   pinned=pinned, h=h, E=E.astype(np.float32), name="ellipsoid")


def read_medit(path: str):
    """Minimal reader for the Medit ``.mesh`` files the reference ships under ``assets/``
    (the reference loads them with meshio in ``utils/io.py:11-19``). Returns
    ``(verts float64 (n,3), tets int32 flat, faces int32 flat)``, ids made 0-based."""
    with open(path, "r") as f:
        toks = " ".join(line.split("#", 1)[0] for line in f).split()
    i = 0
    verts = tets = faces = None
    while i < len(toks):
        key = toks[i]
        if key == "Vertices":
            n = int(toks[i + 1]); i += 2
            arr = np.array(toks[i:i + 4 * n], dtype=np.float64).reshape(n, 4)
            verts = arr[:, :3]
            i += 4 * n
        elif key == "Tetrahedra":
            n = int(toks[i + 1]); i += 2
            arr = np.array(toks[i:i + 5 * n], dtype=np.int64).reshape(n, 5)
            tets = (arr[:, :4] - 1).reshape(-1).astype(np.int32)
            i += 5 * n
        elif key == "Triangles":
            n = int(toks[i + 1]); i += 2
            arr = np.array(toks[i:i + 4 * n], dtype=np.int64).reshape(n, 4)
            faces = (arr[:, :3] - 1).reshape(-1).astype(np.int32)
            i += 4 * n
        elif key == "Edges":
            n = int(toks[i + 1]); i += 2 + 3 * n
        elif key == "Dimension":
            i += 2
        elif key == "MeshVersionFormatted":
            i += 2
        else:
            i += 1
    if verts is None or tets is None:
        raise ValueError(f"{path}: no Vertices/Tetrahedra section")
    return verts, tets, faces


def load_medit_mesh(path: str, scale: float = 1.0, repose=(0.0, 0.0, 0.0),
                    fixed_ids=None, rho: float = 1.0) -> TetMesh:
    """Medit file -> ``TetMesh`` with the reference's scaling/repose and mass convention
    (``data/tet_data.py:24,39-44``); ``fixed_ids`` get ``invm = 0``
    (``test/test_mgpbd_tetmesh.py:43-45``)."""
    verts, tets, _ = read_medit(path)
    verts = verts * scale + np.asarray(repose, dtype=np.float32)   # float64, as TetData
    vm, vol = vertex_mass(verts, tets, rho)
    verts = verts.astype(np.float32)
    invm = (1.0 / vm).astype(np.float32)
    pinned = np.zeros(verts.shape[0], dtype=bool)
    if fixed_ids is not None:
        pinned[np.asarray(fixed_ids, dtype=np.int64)] = True
        invm[pinned] = 0.0
    return TetMesh(verts=verts, tets=tets, invm=invm, tet_mass=vol.astype(np.float32),
                   pinned=pinned, name=path.rsplit("/", 1)[-1])


def drag_schedule(mesh: TetMesh, seed: int, n_frames: int, box: float = 0.5):
    """Spring-drag loading, one (vertex, target) per frame from ``seed``: a uniformly drawn
    free vertex (``invm > 0``) and a target uniform in the axis-aligned cube of side ``box``
    (metres) centred on that vertex's rest position."""
    rng = np.random.default_rng(seed)
    free = np.flatnonzero(mesh.invm > 0)
    out = []
    for _ in range(n_frames):
        v = int(free[rng.integers(len(free))])
        t = mesh.verts[v].astype(np.float64) + rng.uniform(-0.5 * box, 0.5 * box, 3)
        out.append((v, t))
    return out
