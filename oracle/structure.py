"""ORACLE (test infrastructure only) -- literal CPU restatement of the integer/structure part
of the reference MGPBD path. Imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` only, and only as the checker; the product never calls it.

Each function follows one reference function statement by statement, with the same Python
containers (``dict``/``set``/``sorted``) so that CPython's set iteration order -- which
decides the aggregation's tie-breaks -- is the one the reference sees:

* ``build_sparsity_pattern``        cons/mgpbd_gpu.py:153-194
* ``build_tet_face_adjacency``      cons/tet_aggregation.py:11-37
* ``tet_geometry_aggregation``      cons/tet_aggregation.py:40-127
* ``build_prolongation_from_aggregates`` cons/tet_aggregation.py:130-142
* ``build_tet_hierarchy``           cons/tet_aggregation.py:145-211
* ``galerkin_contribution_maps``    cons/mgpcg_gpu.py:43-99
* ``graph_aggregate``               the growth rule of cons/tet_aggregation.py:80-111 on a CSR
                                    graph (the extended hierarchy's coarse levels)
* ``merge_singletons``              extension, no reference counterpart
* ``standard_aggregation``          pyamg's standard aggregation (the aggregation='pyamg'
                                    branch, cons/mgpbd_gpu.py:124-134; pyamg is absent: its
                                    published three-pass algorithm, parity unpinned against
                                    pyamg itself)

Pinning (tests/test_structure.py): the CSR pattern against brute-force ``B B^T != 0``
of the vertex incidence matrix B; face adjacency against ``(B B^T)_ij == 3``; the
aggregation's O(deg) candidate form against the literal O(n) set difference; Galerkin
patterns against scipy ``P^T A P``; contribution maps against a dense re-summation;
``graph_aggregate`` / ``merge_singletons`` against the invariants of their rules (every node
assigned once, aggregates connected and within max_agg, seeds in ascending degree, no
singleton left beside a neighbour, compact ids in old-id order) checked by brute force;
``standard_aggregation`` against the invariants of its three passes (pass-1 aggregates are
whole closed neighbourhoods of pairwise non-adjacent roots, each pass-2 node is adjacent to
its aggregate's root neighbourhood, pass-3 aggregates are connected, isolated nodes -1).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def build_sparsity_pattern(t_i_np: np.ndarray, nt: int):
    """cons/mgpbd_gpu.py:153-194: rows = ascending vertex-sharing neighbours, diagonal last."""
    vert_to_tets = {}
    for t in range(nt):
        for l in range(4):
            v = int(t_i_np[t * 4 + l])
            if v not in vert_to_tets:
                vert_to_tets[v] = set()
            vert_to_tets[v].add(t)
    adj = {}
    for t in range(nt):
        neighbors = set()
        for l in range(4):
            v = int(t_i_np[t * 4 + l])
            neighbors |= vert_to_tets[v]
        neighbors.discard(t)
        adj[t] = sorted(neighbors)
    num_adj = np.array([len(adj[t]) for t in range(nt)], dtype=np.int32)
    nnz = int(np.sum(num_adj) + nt)
    indptr = np.zeros(nt + 1, dtype=np.int32)
    indices = np.zeros(nnz, dtype=np.int32)
    for i in range(nt):
        indptr[i + 1] = indptr[i] + num_adj[i] + 1
        indices[indptr[i]:indptr[i] + num_adj[i]] = adj[i]
        indices[indptr[i] + num_adj[i]] = i
    return indptr, indices


def build_tet_face_adjacency(t_i_np: np.ndarray, n_tets: int):
    """cons/tet_aggregation.py:11-37 (dict of faces in first-seen order; pairs -> sets)."""
    face_to_tets = {}
    for t in range(n_tets):
        verts = [int(t_i_np[t * 4 + l]) for l in range(4)]
        for skip in range(4):
            face = tuple(sorted(verts[j] for j in range(4) if j != skip))
            if face not in face_to_tets:
                face_to_tets[face] = []
            face_to_tets[face].append(t)
    adj = {t: set() for t in range(n_tets)}
    for face, tets in face_to_tets.items():
        if len(tets) == 2:
            adj[tets[0]].add(tets[1])
            adj[tets[1]].add(tets[0])
    return adj


def tet_geometry_aggregation(t_i_np, v_p_np, n_tets, max_agg_size=8, literal=None):
    """cons/tet_aggregation.py:40-127.

    ``literal=True`` evaluates the candidate list exactly as written,
    ``adj[seed] - set(np.where(visited)[0])`` (O(n) per seed); ``literal=False`` uses
    ``{t for t in adj[seed] if not visited[t]}``, which inserts the same elements in the same
    order into a fresh set (CPython iterates the left operand when it is not more than 4x
    larger than the right one, and ``visited[seed]`` is already set), so the resulting
    iteration order is identical. Default: literal for n_tets <= 20000.
    """
    if literal is None:
        literal = n_tets <= 20000
    adj = build_tet_face_adjacency(t_i_np, n_tets)
    agg_ids = -np.ones(n_tets, dtype=np.int32)
    n_agg = 0
    visited = np.zeros(n_tets, dtype=bool)
    order = sorted(range(n_tets), key=lambda t: len(adj[t]))
    for seed in order:
        if visited[seed]:
            continue
        agg_ids[seed] = n_agg
        visited[seed] = True
        current_size = 1
        if literal:
            cand_set = adj[seed] - set(np.where(visited)[0])
        else:
            cand_set = {t for t in adj[seed] if not visited[t]}
        candidates = sorted(cand_set, key=lambda t: len(adj[t]))
        for nb in candidates:
            if visited[nb]:
                continue
            if current_size >= max_agg_size:
                break
            agg_ids[nb] = n_agg
            visited[nb] = True
            current_size += 1
            if current_size < max_agg_size:
                for nb2 in adj[nb]:
                    if not visited[nb2] and current_size < max_agg_size:
                        agg_ids[nb2] = n_agg
                        visited[nb2] = True
                        current_size += 1
        n_agg += 1
    # Pass 2 of the reference (cons/tet_aggregation.py:108-124) assigns tets left at -1 to
    # the nearest aggregate centroid; pass 1 above makes every unvisited tet a seed, so no
    # tet is ever left at -1 and pass 2 never runs.
    assert np.all(agg_ids >= 0)
    return agg_ids, n_agg


def build_prolongation_from_aggregates(agg_ids, n_tets, n_agg):
    """cons/tet_aggregation.py:130-142: P[i, agg[i]] = 1 (FP64 CSR)."""
    rows = np.arange(n_tets, dtype=np.int32)
    cols = agg_ids.astype(np.int32)
    data = np.ones(n_tets, dtype=np.float64)
    return sp.csr_matrix((data, (rows, cols)), shape=(n_tets, n_agg))


def build_tet_hierarchy(t_i_np, v_p_np, n_tets, max_coarse=400, max_levels=4):
    """cons/tet_aggregation.py:145-211, including the virtual-tet coarse geometry whose
    faces are all ``(a, a, a)`` (so no coarse adjacency exists and the second attempt makes
    singletons and is popped)."""
    Ps = []
    sizes = [n_tets]
    current_t_i = np.asarray(t_i_np)
    current_v_p = np.asarray(v_p_np)
    current_ntets = n_tets
    for level in range(max_levels):
        if current_ntets <= max_coarse:
            break
        max_agg = min(8, max(2, current_ntets // max(max_coarse, 1)))
        max_agg = max(max_agg, 2)
        agg_ids, n_agg = tet_geometry_aggregation(current_t_i, current_v_p, current_ntets,
                                                  max_agg_size=max_agg)
        P = build_prolongation_from_aggregates(agg_ids, current_ntets, n_agg)
        Ps.append(P)
        sizes.append(n_agg)
        if n_agg >= current_ntets:
            Ps.pop()
            sizes.pop()
            break
        agg_centroids = np.zeros((n_agg, 3), dtype=np.float64)
        agg_counts = np.zeros(n_agg, dtype=np.float64)
        for t in range(current_ntets):
            c = np.zeros(3, dtype=np.float64)
            for l in range(4):
                c += current_v_p[current_t_i[t * 4 + l]]
            c /= 4.0
            agg_centroids[agg_ids[t]] += c
            agg_counts[agg_ids[t]] += 1.0
        for a in range(n_agg):
            if agg_counts[a] > 0:
                agg_centroids[a] /= agg_counts[a]
        virtual_t_i = np.zeros(n_agg * 4, dtype=np.int32)
        for a in range(n_agg):
            for l in range(4):
                virtual_t_i[a * 4 + l] = a
        current_t_i = virtual_t_i
        current_v_p = agg_centroids
        current_ntets = n_agg
    return Ps, sizes


def galerkin_contribution_maps(A_fine: sp.csr_matrix, P: sp.csr_matrix):
    """cons/mgpcg_gpu.py:46-99 for one level: returns (Ac FP64 CSR, fine_idx, p_ki, p_lj,
    offsets) with the reference's loop order (coarse row, coarse col, fine row of agg i,
    fine CSR entry)."""
    A_fine = A_fine.tocsr()
    Ac = (P.T @ A_fine @ P).tocsr()
    PT = P.T.tocsr()
    offsets = [0]
    fine_idx, p_ki, p_lj = [], [], []
    for i in range(Ac.shape[0]):
        for ptr_c in range(Ac.indptr[i], Ac.indptr[i + 1]):
            j = Ac.indices[ptr_c]
            fine_rows_i = PT.indices[PT.indptr[i]:PT.indptr[i + 1]]
            p_vals_i = PT.data[PT.indptr[i]:PT.indptr[i + 1]]
            fine_rows_j = PT.indices[PT.indptr[j]:PT.indptr[j + 1]]
            p_vals_j = PT.data[PT.indptr[j]:PT.indptr[j + 1]]
            for ki_idx, k in enumerate(fine_rows_i):
                pk_i = p_vals_i[ki_idx]
                for a_ptr in range(A_fine.indptr[k], A_fine.indptr[k + 1]):
                    l = A_fine.indices[a_ptr]
                    lj_pos = np.searchsorted(fine_rows_j, l)
                    if lj_pos < len(fine_rows_j) and fine_rows_j[lj_pos] == l:
                        fine_idx.append(a_ptr)
                        p_ki.append(float(pk_i))
                        p_lj.append(float(p_vals_j[lj_pos]))
            offsets.append(len(fine_idx))
    return (Ac, np.array(fine_idx, dtype=np.int32), np.array(p_ki, dtype=np.float32),
            np.array(p_lj, dtype=np.float32), np.array(offsets, dtype=np.int32))


def graph_aggregate(ptr, idx, max_agg):
    """Greedy growth of the reference's tet aggregation (cons/tet_aggregation.py:80-111:
    seeds in stable ascending-degree order; each seed takes its unvisited neighbours in stable
    ascending-degree order and, while room remains, each taken neighbour's unvisited
    neighbours in adjacency order, up to max_agg) applied to a CSR graph."""
    n = len(ptr) - 1
    deg = [ptr[i + 1] - ptr[i] for i in range(n)]
    order = sorted(range(n), key=lambda t: deg[t])
    visited = [False] * n
    agg = [-1] * n
    n_agg = 0
    for seed in order:
        if visited[seed]:
            continue
        agg[seed] = n_agg
        visited[seed] = True
        size = 1
        cand = sorted([j for j in idx[ptr[seed]:ptr[seed + 1]] if not visited[j]],
                      key=lambda t: deg[t])
        for nb in cand:
            if visited[nb]:
                continue
            if size >= max_agg:
                break
            agg[nb] = n_agg
            visited[nb] = True
            size += 1
            if size < max_agg:
                for nb2 in idx[ptr[nb]:ptr[nb + 1]]:
                    if not visited[nb2] and size < max_agg:
                        agg[nb2] = n_agg
                        visited[nb2] = True
                        size += 1
        n_agg += 1
    return agg, n_agg


def merge_singletons(ptr, idx, agg, n_agg):
    """Extension (no reference counterpart): in node order a singleton joins the smallest
    neighbouring aggregate (ties: lowest id, sizes as updated so far); aggregates renumbered in
    ascending old id, empty ones dropped."""
    agg = list(agg)
    size = [0] * n_agg
    for a in agg:
        size[a] += 1
    for i in range(len(agg)):
        a = agg[i]
        if size[a] != 1:
            continue
        cands = sorted({agg[j] for j in idx[ptr[i]:ptr[i + 1]]} - {a}, key=lambda c: (size[c], c))
        if not cands:
            continue
        size[a] = 0
        size[cands[0]] += 1
        agg[i] = cands[0]
    keep = [a for a in range(n_agg) if size[a] > 0]
    renum = {a: k for k, a in enumerate(keep)}
    return [renum[a] for a in agg], len(keep)


def standard_aggregation(ptr, idx):
    """pyamg's three-pass standard aggregation of a CSR graph (Vanek, Mandel, Brezina 1996):
    pass 1, in node order: a node none of whose neighbours is aggregated forms an aggregate
    with all of them (no off-diagonal neighbour: isolated, -1); pass 2, in node order: a free
    node joins the pass-1 aggregate of its first neighbour, in adjacency order, that pass 1
    aggregated; pass 3, in node order: a free node starts an aggregate with its free
    neighbours. Returns (agg, n_agg)."""
    n = len(ptr) - 1
    FREE, ISO = None, -1
    agg = [FREE] * n
    pass1 = [False] * n
    n_agg = 0
    for i in range(n):
        if agg[i] is not FREE:
            continue
        nbrs = [j for j in idx[ptr[i]:ptr[i + 1]] if j != i]
        if not nbrs:
            agg[i] = ISO
        elif all(agg[j] is FREE for j in nbrs):
            for j in [i] + list(idx[ptr[i]:ptr[i + 1]]):
                agg[j] = n_agg
                pass1[j] = True
            n_agg += 1
    for i in range(n):
        if agg[i] is not FREE:
            continue
        for j in idx[ptr[i]:ptr[i + 1]]:
            if pass1[j] and agg[j] is not FREE and agg[j] != ISO:
                agg[i] = agg[j]
                break
    for i in range(n):
        if agg[i] is not FREE:
            continue
        agg[i] = n_agg
        for j in idx[ptr[i]:ptr[i + 1]]:
            if agg[j] is FREE:
                agg[j] = n_agg
        n_agg += 1
    return [int(a) for a in agg], n_agg
