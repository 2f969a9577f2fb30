"""Host side of the factored level 0's 16-bit slot offsets: the per-vertex base is the smallest
tet at the vertex, every tet at a vertex lies within 65535 of it, and a mesh whose tets at some
vertex are farther apart gets no offsets (the passes then read the incidence list)."""

import numpy as np

from controllable_pbd_3d_amd import _lib
from controllable_pbd_3d_amd.mesh import box_mesh
from controllable_pbd_3d_amd.mgpcg_gpu import slot_offset_base


def _incidence(tets, nv):
    lib = _lib.load()
    t = np.ascontiguousarray(tets.reshape(-1), np.int32)
    n = len(t) // 4
    vptr = np.empty(nv + 1, np.int32)
    vinc = np.empty(4 * n, np.int32)
    _lib.check(lib.mgpbd_vertex_incidence(n, nv, _lib.np_ptr(t), _lib.np_ptr(vptr),
                                          _lib.np_ptr(vinc)))
    return vptr, vinc


def test_base_is_the_smallest_tet_and_offsets_fit():
    m = box_mesh(6, 0.05, jitter=0.05, seed=3)
    vptr, vinc = _incidence(m.tets, m.n_verts)
    base = slot_offset_base(vptr, vinc)
    assert base is not None and base.dtype == np.int32
    for v in range(m.n_verts):        # brute force: every vertex's tets
        tets = np.nonzero((m.tets.reshape(-1, 4) == v).any(axis=1))[0]
        assert base[v] == tets.min()
        assert tets.max() - base[v] <= 65535


def test_far_apart_tets_get_no_offsets():
    # two copies of a block: vertex 0 is shared by tet 0 and by a tet 70000 places later
    m = box_mesh(3, 0.05)
    n = m.n_tets
    reps = 70000 // n + 2
    t4 = m.tets.reshape(-1, 4)
    tets = np.concatenate([t4 + m.n_verts * k for k in range(reps)])
    tets[-1] = t4[0]                  # the last tet reuses the first tet's vertices
    nv = m.n_verts * reps
    vptr, vinc = _incidence(tets, nv)
    assert slot_offset_base(vptr, vinc) is None
    ok = slot_offset_base(*_incidence(tets[:-1], nv))
    assert ok is not None
