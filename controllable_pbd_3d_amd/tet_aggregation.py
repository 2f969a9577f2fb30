"""Tet-aggregation multigrid setup (host side), same interface as the reference's
``cons/tet_aggregation.py``; the work runs in libmgpbd's C++ setup code in O(n).

* ``build_tet_face_adjacency``          cons/tet_aggregation.py:11-37
* ``tet_geometry_aggregation``          cons/tet_aggregation.py:40-127
* ``build_prolongation_from_aggregates`` cons/tet_aggregation.py:130-142
* ``build_tet_hierarchy``               cons/tet_aggregation.py:145-211

The reference hierarchy has at most two levels: its coarse "virtual tets" repeat one id four
times, so they have no shared faces, aggregate into singletons and the attempt is popped.
``build_extended_hierarchy`` (no reference counterpart) keeps coarsening on the aggregate
graph induced by face adjacency, with the same greedy growth rule.

``standard_aggregation`` / ``build_pyamg_hierarchy`` are the ``aggregation='pyamg'`` branch
(cons/mgpbd_gpu.py:124-134): pyamg's standard aggregation on the symmetric strength graph with
theta = 0 (every stored entry of the level matrix), level after level on the Galerkin
patterns until at most ``max_coarse`` unknowns remain. pyamg's prolongator there
(``smooth=None``) is the tentative one, 1/sqrt(|aggregate|) per row; it differs from the unit
P of these maps by a diagonal column scaling, which leaves the V-cycle (Jacobi smoothing,
Galerkin coarse levels, coarse correction) unchanged in exact arithmetic.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import I32P, check, np_ptr


def face_adjacency_arrays(t_i_np, n_tets):
    """(adj int32[4n] in the reference's set iteration order, -1 padded; deg int32[n])."""
    lib = _lib.load()
    t = np.ascontiguousarray(t_i_np, dtype=np.int32).reshape(-1)
    adj = np.empty(4 * n_tets, np.int32)
    deg = np.empty(n_tets, np.int32)
    check(lib.mgpbd_face_adjacency(n_tets, np_ptr(t), np_ptr(adj), np_ptr(deg)),
          "mgpbd_face_adjacency")
    return adj, deg


def build_tet_face_adjacency(t_i_np, n_tets):
    """dict tet -> set of face-adjacent tets (cons/tet_aggregation.py:11-37)."""
    adj, deg = face_adjacency_arrays(t_i_np, n_tets)
    return {t: set(int(v) for v in adj[4 * t:4 * t + deg[t]]) for t in range(n_tets)}


def tet_geometry_aggregation(t_i_np, v_p_np, n_tets, max_agg_size=8):
    """(agg_ids int32[n], n_agg) -- cons/tet_aggregation.py:40-127. ``v_p_np`` is unused, as
    in the reference (only its dead second pass reads centroids)."""
    lib = _lib.load()
    adj, deg = face_adjacency_arrays(t_i_np, n_tets)
    agg = np.empty(n_tets, np.int32)
    n_agg = np.zeros(1, np.int32)
    check(lib.mgpbd_tet_aggregate(n_tets, np_ptr(adj), np_ptr(deg), int(max_agg_size),
                                  np_ptr(agg), np_ptr(n_agg)), "mgpbd_tet_aggregate")
    return agg, int(n_agg[0])


def build_prolongation_from_aggregates(agg_ids, n_tets, n_agg):
    """P[i, agg[i]] = 1 (cons/tet_aggregation.py:130-142)."""
    rows = np.arange(n_tets, dtype=np.int32)
    return sp.csr_matrix((np.ones(n_tets), (rows, np.asarray(agg_ids, np.int32))),
                         shape=(n_tets, n_agg))


def reference_aggregates(t_i_np, n_tets, max_coarse=400, max_levels=4):
    """Aggregate maps of the reference hierarchy: [] or [agg0] (see module docstring)."""
    if max_levels < 1 or n_tets <= max_coarse:
        return []
    max_agg = max(min(8, max(2, n_tets // max(max_coarse, 1))), 2)
    agg, n_agg = tet_geometry_aggregation(t_i_np, None, n_tets, max_agg)
    if n_agg >= n_tets:
        return []
    return [(agg, n_agg)]


def build_tet_hierarchy(t_i_np, v_p_np, n_tets, max_coarse=400, max_levels=4):
    """(Ps, sizes) exactly as cons/tet_aggregation.py:145-211 returns them."""
    aggs = reference_aggregates(t_i_np, n_tets, max_coarse, max_levels)
    Ps = [build_prolongation_from_aggregates(a, n_tets, n) for a, n in aggs]
    sizes = [n_tets] + [n for _, n in aggs]
    return Ps, sizes


def _graph_from_face_adjacency(adj, deg, n):
    """Face adjacency as sorted CSR (ptr, idx)."""
    a = adj.reshape(n, 4).astype(np.int64)
    valid = np.arange(4)[None, :] < deg[:, None]
    a = np.where(valid, a, np.iinfo(np.int64).max)
    a.sort(axis=1)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(deg)
    return ptr, a[valid].astype(np.int32)


def build_extended_hierarchy(t_i_np, n_tets, max_coarse=512, max_levels=8, max_agg=8,
                             first_level="reference", merge_singletons=True):
    """Aggregate maps [(agg_l, n_{l+1})] for a deeper hierarchy (extension, no reference
    counterpart). Level 0 uses the reference's tet aggregation; every further level
    aggregates the graph of aggregates joined by a shared face, with the reference's greedy
    growth rule followed (``merge_singletons``) by merging every singleton into the smallest
    neighbouring aggregate, until at most ``max_coarse`` unknowns remain."""
    lib = _lib.load()
    adj, deg = face_adjacency_arrays(t_i_np, n_tets)
    out = []
    if n_tets <= max_coarse or max_levels < 1:
        return out
    if first_level == "reference":
        ma = max(min(8, max(2, n_tets // 400)), 2)
    else:
        ma = max_agg
    agg = np.empty(n_tets, np.int32)
    n_agg = np.zeros(1, np.int32)
    check(lib.mgpbd_tet_aggregate(n_tets, np_ptr(adj), np_ptr(deg), ma, np_ptr(agg),
                                  np_ptr(n_agg)), "mgpbd_tet_aggregate")
    if int(n_agg[0]) >= n_tets:
        return out
    out.append((agg, int(n_agg[0])))
    gptr, gidx = _graph_from_face_adjacency(adj, deg, n_tets)
    n = n_tets
    while len(out) < max_levels and out[-1][1] > max_coarse:
        a, nc = out[-1]
        g = lib.mgpbd_graph_coarsen(n, np_ptr(gptr), np_ptr(gidx), np_ptr(a), nc)
        if not g:
            check(-1, "mgpbd_graph_coarsen")
        nnz = lib.mgpbd_graph_nnz(g)
        cptr = np.empty(nc + 1, np.int32)
        cidx = np.empty(max(nnz, 1), np.int32)
        check(lib.mgpbd_graph_copy(g, np_ptr(cptr), np_ptr(cidx)), "mgpbd_graph_copy")
        lib.mgpbd_graph_free(g)
        cidx = cidx[:nnz]
        agg2 = np.empty(nc, np.int32)
        check(lib.mgpbd_graph_aggregate(nc, np_ptr(cptr), np_ptr(cidx), int(max_agg),
                                        np_ptr(agg2), np_ptr(n_agg)), "mgpbd_graph_aggregate")
        if merge_singletons:
            check(lib.mgpbd_graph_merge_singletons(nc, np_ptr(cptr), np_ptr(cidx), np_ptr(agg2),
                                                   np_ptr(n_agg)), "mgpbd_graph_merge_singletons")
        if int(n_agg[0]) >= nc:
            break
        out.append((agg2, int(n_agg[0])))
        gptr, gidx, n = cptr, cidx, nc
    return out


def standard_aggregation(indptr, indices):
    """pyamg's standard aggregation of a CSR graph: (agg int32[n] with -1 for isolated nodes,
    n_agg)."""
    lib = _lib.load()
    indptr = np.ascontiguousarray(indptr, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    n = len(indptr) - 1
    agg = np.empty(max(n, 1), np.int32)
    na = np.zeros(1, np.int32)
    check(lib.mgpbd_standard_aggregate(n, np_ptr(indptr), np_ptr(indices), np_ptr(agg),
                                       np_ptr(na)), "mgpbd_standard_aggregate")
    return agg[:n], int(na[0])


def build_pyamg_hierarchy(indptr, indices, values=None, max_coarse=400, max_levels=10):
    """Aggregate maps [(agg_l, n_{l+1})] of the ``aggregation='pyamg'`` hierarchy for a level-0
    matrix pattern. ``values`` (the level-0 values, optional) makes the coarse patterns drop
    entries whose Galerkin sums are exactly zero, as pyamg's sparse products do; without them
    the coarse patterns are structural. Isolated nodes (no off-diagonal entry), which pyamg
    leaves out of every aggregate, become singleton aggregates here (the unit-P transfer needs
    one aggregate per row); coarsening stops when a level does not shrink."""
    import ctypes
    lib = _lib.load()
    ptr = np.ascontiguousarray(indptr, np.int32)
    idx = np.ascontiguousarray(indices, np.int32)
    vals = None if values is None else np.ascontiguousarray(values, np.float32)
    n = len(ptr) - 1
    out = []
    while len(out) + 1 < max_levels and n > max_coarse:
        agg, na = standard_aggregation(ptr, idx)
        iso = np.flatnonzero(agg < 0)
        if len(iso):
            agg[iso] = na + np.arange(len(iso), dtype=np.int32)
            na += len(iso)
        if na >= n or na == 0:
            break
        out.append((agg, na))
        g = lib.mgpbd_galerkin_build(n, np_ptr(ptr), np_ptr(idx),
                                     None if vals is None else vals.ctypes.data_as(_lib.F32P),
                                     np_ptr(agg), na)
        if not g:
            check(-1, "mgpbd_galerkin_build")
        cnnz, ncon = ctypes.c_int64(), ctypes.c_int64()
        check(lib.mgpbd_galerkin_sizes(g, ctypes.byref(cnnz), ctypes.byref(ncon)),
              "mgpbd_galerkin_sizes")
        cptr = np.empty(na + 1, np.int32)
        cidx = np.empty(max(cnnz.value, 1), np.int32)
        cdiag = np.empty(na, np.int32)
        fidx = np.empty(max(ncon.value, 1), np.int32)
        coff = np.empty(cnnz.value + 1, np.int32)
        check(lib.mgpbd_galerkin_copy(g, np_ptr(cptr), np_ptr(cidx), np_ptr(cdiag), np_ptr(fidx),
                                      np_ptr(coff)), "mgpbd_galerkin_copy")
        lib.mgpbd_galerkin_free(g)
        if vals is not None:   # coarse values: FP64 sums of the fine values in map order
            seg = np.repeat(np.arange(cnnz.value), np.diff(coff))
            vals = np.bincount(seg, weights=vals[fidx[:ncon.value]].astype(np.float64),
                               minlength=cnnz.value).astype(np.float32)
        ptr, idx, n = cptr, cidx[:cnnz.value], na
    return out
