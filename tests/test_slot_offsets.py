"""Host side of the factored level 0's 16-bit slot offsets: the per-vertex base is the smallest
tet at the vertex, every tet at a vertex lies within 65535 of it, and a mesh whose tets at some
vertex are farther apart gets no offsets (the passes then read the incidence list)."""

import numpy as np

from controllable_pbd_3d_amd import _lib
from controllable_pbd_3d_amd.mesh import box_mesh
from controllable_pbd_3d_amd.mgpcg_gpu import (packed_rows, slot_offset_base, uniform_rows,
                                              vertex_slices)


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


def test_uniform_layout():
    """A Kuhn block has at most 24 tets per vertex: six rows of four slots per vertex, every
    incidence at its computed position, the rest zero; a mesh with one busy vertex keeps the
    per-slice layout."""
    m = box_mesh(7, 0.05, jitter=0.05, seed=5)
    vptr, vinc = _incidence(m.tets, m.n_verts)
    rows = uniform_rows(vptr)
    assert rows == 6
    vs_ptr, vs_cnt, vs_inc = vertex_slices(vptr, vinc, rows)
    assert np.array_equal(vs_ptr, 64 * rows * np.arange(len(vs_ptr)))
    seen = np.zeros(len(vs_inc), bool)
    for v in range(m.n_verts):
        for k in range(vs_cnt[v]):
            p = (v // 16) * 64 * rows + 64 * (k // 4) + 4 * (v % 16) + k % 4
            assert vs_inc[p] == vinc[vptr[v] + k]
            seen[p] = True
    assert np.all(vs_inc[~seen] == 0)
    # one vertex in 40 tets: a uniform layout would be 10 rows tall everywhere
    t4 = m.tets.reshape(-1, 4).copy()
    hub = int(t4[0, 0])
    t4[:40, 0] = hub
    vptr2, _ = _incidence(t4, m.n_verts)
    assert uniform_rows(vptr2) == 0


def _unpack(words, bb, ob):
    """Decodes packed row words with plain integer arithmetic (the kernels use shifts)."""
    w = words.astype(object)
    v0 = np.array([int(x) & ((1 << bb) - 1) for x in w], np.int64)
    out = [v0]
    for k in range(3):
        f = np.array([(int(x) >> (bb + k * ob)) & ((1 << ob) - 1) for x in w], np.int64)
        f = np.where(f >= 1 << (ob - 1), f - (1 << ob), f)
        out.append(v0 + f)
    return np.stack(out, axis=1)


def test_packed_rows_round_trip():
    """Every row's four vertices come back from its packed word, in order, on a block whose
    tets span up to a layer of vertices (every other row reversed: differences of both
    signs)."""
    m = box_mesh(9, 0.1)
    rows = m.tets.reshape(-1, 4).astype(np.int64).copy()
    rows[1::2] = rows[1::2, ::-1]
    pk = packed_rows(rows, m.n_verts)
    assert pk is not None
    words, bb, ob = pk
    assert bb == int(m.n_verts - 1).bit_length() and bb + 3 * ob <= 64
    d = rows[:, 1:] - rows[:, :1]
    assert d.min() < 0 < d.max()
    assert np.array_equal(_unpack(words, bb, ob), rows)


def test_packed_rows_refuse_differences_that_do_not_fit():
    rows = np.array([[0, 1, 2, 3], [5, 0, 1 << 21, 7]], np.int64)
    assert packed_rows(rows, 1 << 22) is None          # 22 base bits leave 14 per difference
    pk = packed_rows(rows[:1], 16)
    assert pk is not None and np.array_equal(_unpack(*pk), rows[:1])
