"""Integer / structure parity on the CPU (bit-exact): the oracle's literal restatement is
pinned by brute force, and the product's C++ setup code must equal the oracle.

Pins:
* CSR pattern == nonzeros of B B^T (B = tet-vertex incidence), rows ascending, diagonal last;
* face adjacency == pairs with (B B^T)_ij == 3 (three shared vertices = one shared face);
* O(deg) candidate sets == the literal O(n) set difference (same CPython iteration order);
* Galerkin pattern == scipy P^T A P (explicit zeros, value dropping included);
* contribution maps == the reference's triple loop; refill sums == P^T A P values.
"""

import os

import numpy as np
import pytest
import scipy.sparse as sp

from controllable_pbd_3d_amd import tet_aggregation as ta
from controllable_pbd_3d_amd import _lib
from controllable_pbd_3d_amd.mesh import box_mesh, load_medit_mesh
from oracle import structure
from oracle.mgpbd_oracle import galerkin_contribution_maps_fast

GOLD = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def incidence(tets, n_verts):
    t = tets.reshape(-1, 4)
    rows = np.repeat(np.arange(t.shape[0]), 4)
    return sp.csr_matrix((np.ones(t.size), (rows, t.reshape(-1))), shape=(t.shape[0], n_verts))


def meshes():
    out = [("box3", box_mesh(3, 0.1, seed=0)), ("box5", box_mesh(5, 0.05, seed=1)),
           ("box4x2x3", box_mesh(4, 0.05, seed=2, ny=2, nz=3))]
    for f in ("simple_sphere", "spot", "cube_tet"):
        out.append((f, load_medit_mesh(os.path.join(GOLD, f + ".mesh"))))
    return out


MESHES = meshes()


def product_pattern(m):
    lib = _lib.load()
    p = lib.mgpbd_pattern_build(m.n_tets, m.n_verts, _lib.np_ptr(m.tets))
    nnz = lib.mgpbd_pattern_nnz(p)
    indptr = np.empty(m.n_tets + 1, np.int32)
    indices = np.empty(nnz, np.int32)
    _lib.check(lib.mgpbd_pattern_copy(p, _lib.np_ptr(indptr), _lib.np_ptr(indices)))
    lib.mgpbd_pattern_free(p)
    return indptr, indices


@pytest.mark.parametrize("name,m", MESHES, ids=[n for n, _ in MESHES])
def test_pattern(name, m):
    indptr, indices = structure.build_sparsity_pattern(m.tets, m.n_tets)
    B = incidence(m.tets, m.n_verts)
    G = (B @ B.T).tocsr()
    G.sort_indices()
    for i in range(m.n_tets):
        row = indices[indptr[i]:indptr[i + 1]]
        assert row[-1] == i
        nb = G.indices[G.indptr[i]:G.indptr[i + 1]]
        assert np.array_equal(row[:-1], nb[nb != i])
    pi, px = product_pattern(m)
    assert np.array_equal(pi, indptr) and np.array_equal(px, indices)


def test_pattern_nnz_of_the_8cube_block():
    m = box_mesh(8, 0.05, seed=0)
    indptr, _ = product_pattern(m)
    assert (m.n_verts, m.n_tets, int(indptr[-1])) == (729, 3072, 174000)
    assert int(np.diff(indptr).max()) == 71


@pytest.mark.parametrize("name,m", MESHES, ids=[n for n, _ in MESHES])
def test_face_adjacency(name, m):
    adj = structure.build_tet_face_adjacency(m.tets, m.n_tets)
    B = incidence(m.tets, m.n_verts)
    G = (B @ B.T).tocoo()
    pairs = {(int(i), int(j)) for i, j, v in zip(G.row, G.col, G.data) if v == 3 and i != j}
    got = {(t, u) for t, s in adj.items() for u in s}
    assert got == pairs
    a, d = ta.face_adjacency_arrays(m.tets, m.n_tets)
    for t in range(m.n_tets):
        assert list(adj[t]) == list(a[4 * t:4 * t + d[t]]), "CPython set iteration order"


@pytest.mark.parametrize("name,m", MESHES, ids=[n for n, _ in MESHES])
@pytest.mark.parametrize("max_agg", [2, 5, 8])
def test_aggregation(name, m, max_agg):
    a_lit, n_lit = structure.tet_geometry_aggregation(m.tets, m.verts, m.n_tets, max_agg,
                                                      literal=True)
    a_fast, n_fast = structure.tet_geometry_aggregation(m.tets, m.verts, m.n_tets, max_agg,
                                                        literal=False)
    assert n_lit == n_fast and np.array_equal(a_lit, a_fast)
    a_p, n_p = ta.tet_geometry_aggregation(m.tets, m.verts, m.n_tets, max_agg)
    assert n_p == n_lit and np.array_equal(a_p, a_lit)
    sizes = np.bincount(a_p)
    assert sizes.min() >= 1 and sizes.max() <= max_agg


@pytest.mark.parametrize("name,m", MESHES, ids=[n for n, _ in MESHES])
def test_hierarchy(name, m):
    Ps, sizes = structure.build_tet_hierarchy(m.tets, m.verts.astype(np.float64), m.n_tets)
    Pp, sizes_p = ta.build_tet_hierarchy(m.tets, m.verts, m.n_tets)
    assert sizes == sizes_p and len(Ps) <= 1
    for P, Q in zip(Ps, Pp):
        assert (P != Q).nnz == 0


def _galerkin_product(indptr, indices, values, agg, n_c):
    lib = _lib.load()
    import ctypes
    v = None if values is None else np.ascontiguousarray(values, np.float32)
    g = lib.mgpbd_galerkin_build(len(indptr) - 1, _lib.np_ptr(indptr), _lib.np_ptr(indices),
                                 None if v is None else v.ctypes.data_as(_lib.F32P),
                                 _lib.np_ptr(agg), n_c)
    assert g
    a, b = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(lib.mgpbd_galerkin_sizes(g, ctypes.byref(a), ctypes.byref(b)))
    ci = np.empty(n_c + 1, np.int32)
    cx = np.empty(max(a.value, 1), np.int32)
    cd = np.empty(n_c, np.int32)
    fi = np.empty(max(b.value, 1), np.int32)
    off = np.empty(a.value + 1, np.int32)
    _lib.check(lib.mgpbd_galerkin_copy(g, _lib.np_ptr(ci), _lib.np_ptr(cx), _lib.np_ptr(cd),
                                       _lib.np_ptr(fi), _lib.np_ptr(off)))
    lib.mgpbd_galerkin_free(g)
    return ci, cx[:a.value], cd, fi[:b.value], off


@pytest.mark.parametrize("case", ["rest", "random", "sparse_zeros"])
def test_galerkin_value_pattern(case):
    m = box_mesh(5, 0.05, seed=4)
    indptr, indices = structure.build_sparsity_pattern(m.tets, m.n_tets)
    rng = np.random.default_rng(7)
    rows = np.repeat(np.arange(m.n_tets), np.diff(indptr))
    if case == "rest":       # A = diag(alpha) with explicit zero off-diagonals (init at rest)
        vals = np.where(indices == rows, rng.uniform(1, 2, len(indices)), 0.0)
    elif case == "random":
        vals = rng.standard_normal(len(indices))
    else:                     # some off-diagonals zero, some pairs that cancel exactly
        vals = rng.standard_normal(len(indices))
        vals[rng.random(len(indices)) < 0.6] = 0.0
        vals[indices == rows] = 4.0
    vals = vals.astype(np.float32)
    A = sp.csr_matrix((vals.astype(np.float64), indices, indptr), shape=(m.n_tets,) * 2)
    Ps, _ = structure.build_tet_hierarchy(m.tets, m.verts, m.n_tets)
    P = Ps[0]
    Ac = (P.T @ A @ P).tocsr()
    agg = P.indices.astype(np.int32)
    ci, cx, cd, fi, off = _galerkin_product(indptr, indices, vals, agg, P.shape[1])
    assert np.array_equal(ci, Ac.indptr) and np.array_equal(cx, Ac.indices)
    if case == "rest":
        assert Ac.nnz == P.shape[1], "the coarse matrix of a rest-state A is diagonal"
    # literal contribution maps (cons/mgpcg_gpu.py:71-92) and the vectorised form
    Ac2, fine_idx, p_ki, p_lj, offsets = structure.galerkin_contribution_maps(A, P)
    assert np.array_equal(fi, fine_idx) and np.array_equal(off, offsets)
    fi2, off2 = galerkin_contribution_maps_fast(A, agg, Ac)
    assert np.array_equal(fi2, fine_idx) and np.array_equal(off2, offsets)
    assert np.all(p_ki == 1.0) and np.all(p_lj == 1.0)
    # refill sums reproduce P^T A P
    sums = np.add.reduceat(vals.astype(np.float64)[fi], off[:-1]) if len(fi) else []
    np.testing.assert_allclose(sums, Ac.data, rtol=1e-12, atol=1e-12)
    # the diagonal position of every coarse row
    for I in range(P.shape[1]):
        row = cx[ci[I]:ci[I + 1]]
        assert cd[I] == (ci[I] + int(np.nonzero(row == I)[0][0]) if np.any(row == I) else -1)


def test_galerkin_structural_pattern():
    m = box_mesh(5, 0.05, seed=4)
    indptr, indices = structure.build_sparsity_pattern(m.tets, m.n_tets)
    S = sp.csr_matrix((np.ones(len(indices)), indices, indptr), shape=(m.n_tets,) * 2)
    Ps, _ = structure.build_tet_hierarchy(m.tets, m.verts, m.n_tets)
    Ac = (Ps[0].T @ S @ Ps[0]).tocsr()
    ci, cx, cd, fi, off = _galerkin_product(indptr, indices, None,
                                            Ps[0].indices.astype(np.int32), Ps[0].shape[1])
    assert np.array_equal(ci, Ac.indptr) and np.array_equal(cx, Ac.indices)
    assert len(fi) == len(indices)   # every fine entry lands in some coarse entry
    assert np.array_equal(np.sort(fi), np.arange(len(indices)))


def test_extended_hierarchy_invariants():
    m = box_mesh(8, 0.05, seed=0)
    aggs = ta.build_extended_hierarchy(m.tets, m.n_tets, max_coarse=32)
    sizes = [m.n_tets] + [n for _, n in aggs]
    assert sizes[-1] <= 32 and all(b < a for a, b in zip(sizes, sizes[1:]))
    ref = ta.reference_aggregates(m.tets, m.n_tets)
    assert np.array_equal(aggs[0][0], ref[0][0])       # level 0 is the reference's
    for (agg, n_c), n_f in zip(aggs, sizes):
        assert agg.shape == (n_f,) and agg.min() == 0 and agg.max() == n_c - 1
        assert np.all(np.bincount(agg, minlength=n_c) >= 1)


def _random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=p, random_state=rng, format="csr")
    M = ((M + M.T) != 0).astype(np.int32).tolil()
    M.setdiag(0)
    M = M.tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32)


def _check_aggregates(ptr, idx, agg, n_agg, max_size=None):
    """Every node in one aggregate, ids compact, aggregates connected within themselves."""
    agg = np.asarray(agg)
    n = len(ptr) - 1
    assert agg.shape == (n,) and (n == 0 or (agg.min() == 0 and agg.max() == n_agg - 1))
    sizes = np.bincount(agg, minlength=n_agg)
    assert np.all(sizes >= 1)
    if max_size is not None:
        assert sizes.max() <= max_size
    for a in range(n_agg):
        members = set(np.flatnonzero(agg == a).tolist())
        start = next(iter(members))
        seen, stack = {start}, [start]
        while stack:
            u = stack.pop()
            for v in idx[ptr[u]:ptr[u + 1]]:
                if v in members and v not in seen:
                    seen.add(int(v))
                    stack.append(int(v))
        assert seen == members, a


@pytest.mark.parametrize("n,p,seed,max_agg", [(60, 0.05, 0, 8), (200, 0.02, 1, 8),
                                             (300, 0.01, 2, 4), (150, 0.1, 3, 8)])
def test_graph_aggregate_and_singleton_merge_match_oracle(n, p, seed, max_agg):
    """libmgpbd's graph aggregation and singleton merge equal the oracle restatements, and
    both satisfy their rules' invariants: aggregates connected and within max_agg after the
    growth, seeds (lowest-degree members) in ascending degree; after the merge aggregates
    are still connected and no singleton has a neighbour."""
    from oracle import structure as st
    lib = _lib.load()
    ptr, idx = _random_graph(n, p, seed)
    agg = np.empty(n, np.int32)
    na = np.zeros(1, np.int32)
    _lib.check(lib.mgpbd_graph_aggregate(n, _lib.np_ptr(ptr), _lib.np_ptr(idx), max_agg,
                                         _lib.np_ptr(agg), _lib.np_ptr(na)))
    ref_agg, ref_n = st.graph_aggregate(ptr.tolist(), idx.tolist(), max_agg)
    assert int(na[0]) == ref_n and agg.tolist() == ref_agg
    _check_aggregates(ptr, idx, agg, int(na[0]), max_agg)
    deg = np.diff(ptr)
    seeds = [min(np.flatnonzero(agg == a), key=lambda t: (deg[t], t)) for a in range(int(na[0]))]
    assert all(deg[s] <= deg[t] for s, t in zip(seeds, seeds[1:]))   # seeds by degree
    merged = agg.copy()
    nm = na.copy()
    _lib.check(lib.mgpbd_graph_merge_singletons(n, _lib.np_ptr(ptr), _lib.np_ptr(idx),
                                                _lib.np_ptr(merged), _lib.np_ptr(nm)))
    ref_m, ref_nm = st.merge_singletons(ptr.tolist(), idx.tolist(), agg.tolist(), int(na[0]))
    assert int(nm[0]) == ref_nm and merged.tolist() == ref_m
    _check_aggregates(ptr, idx, merged, int(nm[0]))
    sizes = np.bincount(merged, minlength=int(nm[0]))
    for i in range(n):
        if sizes[merged[i]] == 1:
            assert ptr[i + 1] == ptr[i]          # only isolated nodes stay alone


def test_extended_hierarchy_coarsens_an_irregular_mesh():
    """On a clipped (ellipsoidal) grid the graph levels keep coarsening by at least 2x per
    level down to max_coarse, which the greedy growth alone does not."""
    from controllable_pbd_3d_amd.mesh import ellipsoid_mesh
    m = ellipsoid_mesh(0.04, seed=0)
    aggs = ta.build_extended_hierarchy(m.tets, m.n_tets, max_coarse=64, max_levels=8)
    sizes = [m.n_tets] + [n for _, n in aggs]
    plain = ta.build_extended_hierarchy(m.tets, m.n_tets, max_coarse=64, max_levels=8,
                                        merge_singletons=False)
    print(f"[ellipsoid h=0.04] levels {sizes}; without the merge "
          f"{[m.n_tets] + [n for _, n in plain]}")
    assert sizes[-1] <= 64
    assert all(2 * b <= a for a, b in zip(sizes[1:], sizes[2:]))


@pytest.mark.parametrize("n,p,seed", [(80, 0.05, 0), (300, 0.02, 1), (200, 0.004, 2),
                                      (150, 0.1, 3)])
def test_standard_aggregation_matches_oracle(n, p, seed):
    """pyamg's standard aggregation (the aggregation='pyamg' branch): libmgpbd == the oracle's
    three-pass restatement on random graphs with isolated nodes; every non-isolated node in
    exactly one connected aggregate, isolated nodes (no off-diagonal neighbour) -1."""
    from oracle import structure as st
    ptr, idx = _random_graph(n, p, seed)
    agg, na = ta.standard_aggregation(ptr, idx)
    ref, ref_n = st.standard_aggregation(ptr.tolist(), idx.tolist())
    assert na == ref_n and agg.tolist() == ref
    deg = np.diff(ptr)
    assert np.array_equal(agg < 0, deg == 0)
    if na:
        for a in range(na):
            members = set(np.flatnonzero(agg == a).tolist())
            assert members
            start = next(iter(members))
            seen, stack = {start}, [start]
            while stack:
                u = stack.pop()
                for v in idx[ptr[u]:ptr[u + 1]]:
                    if v in members and v not in seen:
                        seen.add(int(v))
                        stack.append(int(v))
            assert seen == members


def test_pyamg_hierarchy_on_a_block():
    """The aggregation='pyamg' hierarchy on a block's matrix pattern: each level is pyamg's
    standard aggregation of the previous level's Galerkin pattern, sizes shrink to <= 400;
    with the values of A at rest (zero off-diagonals) the first coarse pattern is diagonal,
    every coarse node is isolated and coarsening stops after one level."""
    m = box_mesh(8, 0.05, seed=0)
    indptr, indices = product_pattern(m)
    aggs = ta.build_pyamg_hierarchy(indptr, indices, None, max_coarse=400)
    sizes = [m.n_tets] + [n for _, n in aggs]
    print(f"[pyamg hierarchy] {sizes}")
    assert sizes[-1] <= 400 and all(b < a for a, b in zip(sizes, sizes[1:]))
    agg0, n1 = ta.standard_aggregation(indptr, indices)
    assert np.array_equal(aggs[0][0], agg0) and aggs[0][1] == n1
    vals = np.where(indices == np.repeat(np.arange(m.n_tets), np.diff(indptr)), 1.0, 0.0)
    at_rest = ta.build_pyamg_hierarchy(indptr, indices, vals.astype(np.float32), max_coarse=400)
    assert len(at_rest) == 1 and np.array_equal(at_rest[0][0], agg0)


def test_column_scaled_prolongation_gives_the_same_vcycle():
    """pyamg's tentative prolongator (1/sqrt(|aggregate|) per row) and the unit P of the same
    aggregates give the same two-level V-cycle (weighted Jacobi, Galerkin coarse level,
    exact coarse solve) up to rounding: a column scaling S gives A_c' = S A_c S and
    P' A_c'^-1 P'^T = P A_c^-1 P^T."""
    rng = np.random.default_rng(0)
    n = 120
    M = sp.random(n, n, density=0.06, random_state=rng)
    A = (M @ M.T + sp.eye(n) * 0.5).tocsr()
    agg, na = ta.standard_aggregation(A.indptr, A.indices)
    agg = np.where(agg < 0, na + np.arange(n), agg)
    _, agg = np.unique(agg, return_inverse=True)
    nc = agg.max() + 1
    P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
    size = np.bincount(agg)
    Ps = sp.csr_matrix((1 / np.sqrt(size[agg]), (np.arange(n), agg)), shape=(n, nc))
    d = A.diagonal()

    def vcycle(P, b):
        x = np.zeros(n)
        for _ in range(2):
            x = x + (2 / 3) * (b - A @ x) / d
        Ac = (P.T @ A @ P).toarray()
        x = x + P @ np.linalg.solve(Ac, P.T @ (b - A @ x))
        for _ in range(2):
            x = x + (2 / 3) * (b - A @ x) / d
        return x

    b = rng.standard_normal(n)
    x1, x2 = vcycle(P, b), vcycle(Ps, b)
    assert np.abs(x1 - x2).max() <= 1e-10 * np.abs(x1).max()
    from controllable_pbd_3d_amd.mgpcg_gpu import _aggs_from_Ps
    assert np.array_equal(_aggs_from_Ps([Ps])[0][0], agg.astype(np.int32))


def test_sell_layout_roundtrip():
    """SELL-64x4: width a multiple of 4 covering the longest row of the slice; entry k of row
    r at slice_ptr[r/64] + 256 (k/4) + 4 (r%64) + k%4; padding points at the row itself."""
    from controllable_pbd_3d_amd.sparse_gpu import sell_layout
    m = box_mesh(4, 0.05, seed=0)
    indptr, indices = product_pattern(m)
    n = m.n_tets
    slice_ptr, sell_idx, pos = sell_layout(indptr, indices, n)
    assert len(np.unique(pos)) == len(pos)
    assert np.array_equal(sell_idx[pos], indices)
    lens = np.diff(indptr)
    for s in range((n + 63) // 64):
        w = (slice_ptr[s + 1] - slice_ptr[s]) // 64
        assert w % 4 == 0 and w - 4 < lens[64 * s:64 * s + 64].max() <= w
        for l in range(64):
            r = 64 * s + l
            for k in range(w):
                p = slice_ptr[s] + 256 * (k // 4) + 4 * l + k % 4
                c = sell_idx[p]
                if r >= n:
                    assert c == 0
                elif k >= lens[r]:
                    assert c == r
                else:
                    assert p == pos[indptr[r] + k] and c == indices[indptr[r] + k]


def test_vertex_incidence():
    m = box_mesh(3, 0.1, seed=0)
    lib = _lib.load()
    vptr = np.empty(m.n_verts + 1, np.int32)
    vinc = np.empty(4 * m.n_tets, np.int32)
    _lib.check(lib.mgpbd_vertex_incidence(m.n_tets, m.n_verts, _lib.np_ptr(m.tets),
                                          _lib.np_ptr(vptr), _lib.np_ptr(vinc)))
    for v in range(m.n_verts):
        e = vinc[vptr[v]:vptr[v + 1]]
        assert np.all(np.diff(e) > 0)
        assert np.all(m.tets[e] == v)
    assert vptr[-1] == 4 * m.n_tets


def test_oracle_c_pattern_equals_literal():
    from oracle import mgpbd_oracle as om
    import ctypes
    for name, m in MESHES:
        L, _, _ = om.lib("f32")
        indptr = np.zeros(m.n_tets + 1, np.int32)
        nnz = L.orc_pattern(m.n_tets, m.n_verts, om.ptr(m.tets, om.I32P),
                            om.ptr(indptr, om.I32P), om.ptr(None, om.I32P))
        indices = np.zeros(nnz, np.int32)
        L.orc_pattern(m.n_tets, m.n_verts, om.ptr(m.tets, om.I32P), om.ptr(indptr, om.I32P),
                      om.ptr(indices, om.I32P))
        ip, ix = structure.build_sparsity_pattern(m.tets, m.n_tets)
        assert np.array_equal(indptr, ip) and np.array_equal(indices, ix), name


@pytest.mark.parametrize("n_rows,seed", [(1, 0), (255, 1), (700, 2), (1000, 3)])
def test_sell_window_lists_each_blocks_segments(n_rows, seed):
    """LDS windows: for each block of 256 rows (over the slice-padded row range), the sorted
    distinct 32-column segments of its rows' columns and own indices; checked against a set
    computed row by row, on random sparse patterns with columns up to 4x the row count and
    empty rows."""
    from controllable_pbd_3d_amd import _lib
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n_rows)
    lens[rng.random(n_rows) < 0.1] = 0
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    indices = rng.integers(0, 4 * n_rows, int(indptr[-1])).astype(np.int32)
    lib = _lib.load()
    nb = (64 * ((n_rows + 63) // 64) + 255) // 256
    seg_ptr = np.zeros(nb + 1, np.int32)
    _lib.check(lib.mgpbd_sell_window(n_rows, _lib.np_ptr(indptr), _lib.np_ptr(indices),
                                     _lib.np_ptr(seg_ptr), None, 0), "mgpbd_sell_window")
    segs = np.zeros(max(int(seg_ptr[-1]), 1), np.int32)
    assert lib.mgpbd_sell_window(n_rows, _lib.np_ptr(indptr), _lib.np_ptr(indices),
                                 _lib.np_ptr(seg_ptr), _lib.np_ptr(segs), len(segs) - 1) == -1
    _lib.check(lib.mgpbd_sell_window(n_rows, _lib.np_ptr(indptr), _lib.np_ptr(indices),
                                     _lib.np_ptr(seg_ptr), _lib.np_ptr(segs), len(segs)),
               "mgpbd_sell_window")
    for b in range(nb):
        want = set()
        for r in range(256 * b, min(n_rows, 256 * b + 256)):
            want.add(r // 32)
            want.update((indices[indptr[r]:indptr[r + 1]] // 32).tolist())
        got = segs[seg_ptr[b]:seg_ptr[b + 1]]
        assert got.tolist() == sorted(want), b
